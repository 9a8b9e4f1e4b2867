! shear_rk3.f90 -- Fortran 2003 bind(C) interface of libshear_rk3 (the C ABI
! declared in shear_rk3.h), for the reference driver: `use snf_gpu` where
! `call cross(...)` was (main.f90:279, gmres_shear.f90:3103,3370,
! shoot_shear.f90:527) and link with -lshear_rk3.
!
! The derived types mirror struct snf_params and struct snf_run_status field
! by field (same order, same kinds), so c_sizeof of each equals the C sizeof.
! Arrays are passed as assumed-size real(c_double) in the layouts shear_rk3.h
! documents (vor, phi in the reference's jik layout for snf_state_upload /
! snf_state_download).
module snf_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SNF_OK = 0, SNF_BAD_PARAM = 1, SNF_DEVICE_ERROR = 2, &
       SNF_NO_DEVICE = 3, SNF_NOT_READY = 4
  integer(c_int32_t), parameter :: SNF_RUNNING = 0, SNF_STOP = 1, SNF_DUMP = 2
  integer(c_int32_t), parameter :: SNF_STOP_NONE = 0, SNF_STOP_ETIME = 1, SNF_STOP_DT = 2, &
       SNF_STOP_NAN = 3, SNF_STOP_LAMINAR = 4
  integer(c_int32_t), parameter :: SNF_PREC_FP64 = 0, SNF_PREC_MIXED = 1
  integer(c_int32_t), parameter :: SNF_UNER_LEN = 24
  integer(c_int32_t), parameter :: SNF_K_NCLASS = 5

  type, bind(C) :: snf_params
     integer(c_int32_t) :: mgalx, my, mgalz
     real(c_double)     :: Re, alp, gam, Ly, s, chi, CFL
     integer(c_int32_t) :: explicit_visc, nhist, readflag
     real(c_double)     :: uprim, vbulk, etime, dumptime, dtimag
     integer(c_int32_t) :: precision, device, ydealias
  end type snf_params

  type, bind(C) :: snf_run_status
     integer(c_int32_t) :: steps_done, stop
     real(c_double)     :: time, Deltat, xoffb, xofft, umax(0:3), CFLv
     integer(c_int32_t) :: nhist_records, stop_reason
  end type snf_run_status

  interface
     ! ---- lifetime
     integer(c_int) function snf_create(p, h) bind(C, name="snf_create")
       import :: c_int, c_ptr, snf_params
       type(snf_params), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function snf_create
     integer(c_int) function snf_destroy(h) bind(C, name="snf_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function snf_destroy
     type(c_ptr) function snf_last_error() bind(C, name="snf_last_error")
       import :: c_ptr
     end function snf_last_error
     integer(c_int) function snf_sizes(h, kxn, mz, nc) bind(C, name="snf_sizes")
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: kxn, mz
       integer(c_int64_t), intent(out) :: nc
     end function snf_sizes
     ! ---- state transfer
     integer(c_int) function snf_state_upload(h, vor, phi, u00, w00, time, xoffb, xofft) &
          bind(C, name="snf_state_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: vor(*), phi(*), u00(*), w00(*)
       real(c_double), value :: time, xoffb, xofft
     end function snf_state_upload
     integer(c_int) function snf_state_download(h, vor, phi, u00, w00, time, xoffb, xofft) &
          bind(C, name="snf_state_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: vor(*), phi(*), u00(*), w00(*)
       real(c_double), intent(out) :: time, xoffb, xofft
     end function snf_state_download
     integer(c_int) function snf_state_upload_ykxkz(h, vor, phi, u00, w00, time, xoffb, xofft) &
          bind(C, name="snf_state_upload_ykxkz")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: vor(*), phi(*), u00(*), w00(*)
       real(c_double), value :: time, xoffb, xofft
     end function snf_state_upload_ykxkz
     integer(c_int) function snf_state_download_ykxkz(h, vor, phi, u00, w00) &
          bind(C, name="snf_state_download_ykxkz")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: vor(*), phi(*), u00(*), w00(*)
     end function snf_state_download_ykxkz
     ! ---- the time advance (replaces call cross, cross.f90:3-1101)
     integer(c_int) function snf_advance(h, nstep, st) bind(C, name="snf_advance")
       import :: c_int, c_ptr, c_int32_t, snf_run_status
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nstep
       type(snf_run_status), intent(out) :: st
     end function snf_advance
     integer(c_int) function snf_advance_batch(hs, n, nstep, st) bind(C, name="snf_advance_batch")
       import :: c_int, c_ptr, c_int32_t, snf_run_status
       type(c_ptr), intent(in) :: hs(*)
       integer(c_int32_t), value :: n, nstep
       type(snf_run_status), intent(out) :: st(*)
     end function snf_advance_batch
     ! ---- shooting (calc_dns_disk3d, matvec_dns_disk3d)
     integer(c_int) function snf_state_len(h, n) bind(C, name="snf_state_len")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int64_t), intent(out) :: n
     end function snf_state_len
     integer(c_int) function snf_shoot(h, x, t0, xoffb0, xofft0, T, y, st) bind(C, name="snf_shoot")
       import :: c_int, c_ptr, c_double, snf_run_status
       type(c_ptr), value :: h
       real(c_double), intent(in) :: x(*)
       real(c_double), value :: t0, xoffb0, xofft0, T
       real(c_double), intent(out) :: y(*)
       type(snf_run_status), intent(out) :: st
     end function snf_shoot
     integer(c_int) function snf_jvp_batch(h, x0, y0, dirs, ndir, eps, t0, xoffb0, xofft0, T, out) &
          bind(C, name="snf_jvp_batch")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: x0(*), y0(*), dirs(*), eps(*)
       integer(c_int32_t), value :: ndir
       real(c_double), value :: t0, xoffb0, xofft0, T
       real(c_double), intent(out) :: out(*)
     end function snf_jvp_batch
     ! ---- statistics (hist, cross.f90:1773-2121; escru, readwrite.f90:1164-1334)
     integer(c_int) function snf_history(h, out, max_records, n) bind(C, name="snf_history")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(24, *)
       integer(c_int32_t), value :: max_records
       integer(c_int32_t), intent(out) :: n
     end function snf_history
     integer(c_int) function snf_statistics(h, stats, stats2, nacum, total_dt) &
          bind(C, name="snf_statistics")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: stats(23, *), stats2(23, *)
       integer(c_int32_t), intent(out) :: nacum
       real(c_double), intent(out) :: total_dt
     end function snf_statistics
     integer(c_int) function snf_spectra(h, sp, sp2) bind(C, name="snf_spectra")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: sp(*), sp2(*)
     end function snf_spectra
     integer(c_int) function snf_reset_statistics(h) bind(C, name="snf_reset_statistics")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function snf_reset_statistics
     ! ---- operators (derivyc, lapcdy, getv, energy, fourxz, hvhg)
     integer(c_int) function snf_op_derivy(h, u, du, nl, shp, shm, order) bind(C, name="snf_op_derivy")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), shp(*), shm(*)
       real(c_double), intent(out) :: du(*)
       integer(c_int32_t), value :: nl, order
     end function snf_op_derivy
     integer(c_int) function snf_op_helmholtz(h, u, p, nl, k2, shp, shm) bind(C, name="snf_op_helmholtz")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), k2(*), shp(*), shm(*)
       real(c_double), intent(out) :: p(*)
       integer(c_int32_t), value :: nl
     end function snf_op_helmholtz
     integer(c_int) function snf_op_getv(h, fin, fout, dir) bind(C, name="snf_op_getv")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: fin(*)
       real(c_double), intent(inout) :: fout(*)
       integer(c_int32_t), value :: dir
     end function snf_op_getv
     integer(c_int) function snf_energy(h, energ) bind(C, name="snf_energy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: energ(4)
     end function snf_energy
     ! no counterpart in the reference: snf_energy allocating no device memory
     integer(c_int) function snf_energy_noalloc(h, energ) bind(C, name="snf_energy_noalloc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: energ(4)
     end function snf_energy_noalloc
     integer(c_int) function snf_op_fourxz(h, fin, fout, nplanes, dir) bind(C, name="snf_op_fourxz")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: fin(*)
       real(c_double), intent(out) :: fout(*)
       integer(c_int32_t), value :: nplanes, dir
     end function snf_op_fourxz
     integer(c_int) function snf_op_hvhg(h, hg_v, hv_dv, dvordy, out_dvordy, out_hg, out_hv, &
          rf0u, rf0w) bind(C, name="snf_op_hvhg")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: hg_v(*), hv_dv(*), dvordy(*)
       real(c_double), intent(out) :: out_dvordy(*), out_hg(*), out_hv(*), rf0u(*), rf0w(*)
     end function snf_op_hvhg
     ! ---- measurement
     integer(c_int) function snf_set_timing(h, on) bind(C, name="snf_set_timing")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: on
     end function snf_set_timing
     integer(c_int) function snf_get_timing(h, ms, counts, nclass) bind(C, name="snf_get_timing")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: ms(*)
       integer(c_int32_t), intent(out) :: counts(*)
       integer(c_int32_t), value :: nclass
     end function snf_get_timing
     integer(c_int) function snf_copy_probe(device, nbytes, iters, ms) bind(C, name="snf_copy_probe")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int32_t), value :: device, iters
       integer(c_int64_t), value :: nbytes
       real(c_double), intent(out) :: ms
     end function snf_copy_probe
  end interface
end module snf_gpu
