/*
 * shear_rk3.h -- C ABI of the MI355X (gfx950) RK3 time advance for the
 * shear-periodic box (drop-in for `cross` of hiiragiyua/shear_nnf).
 *
 * The reference entry point is the Fortran subroutine
 *   cross(vor,phi,u00,w00,hv,hg,phiwk,vorwk,dvordy,rhvc1,rhvc2,rhgc1,rhgc2,chwk)
 *   (cross.f90:3-4), whose grid, run-control and boundary state live in the
 *   modules ctes / running / bcs (mod.f90:1-130).  Here that state is a
 *   POD parameter block given once at creation, a device-resident handle,
 *   and explicit upload / advance / download calls.  Every function returns
 *   an snf_status code and never exits the process.
 *
 * Host arrays in the reference layout ("jik", cross.f90:32-33):
 *   vor, phi : complex(0:my-1, 0:mx/2-1, 0:mz-1) stored as interleaved
 *              real/imag doubles, y fastest, then kx, then kz
 *   u00, w00 : real(0:my-1)
 * Device storage is kz-chunk-major, [kx][kz/4][y][kz%4]; the handle
 * transposes on upload/download (on the device).  All arithmetic is IEEE
 * binary64 unless the handle was created with SNF_PREC_MIXED.
 *
 * The library exports these functions only (built with hidden visibility).
 */
#ifndef SHEAR_RK3_H
#define SHEAR_RK3_H

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define SNF_API __attribute__((visibility("default")))
#else
#define SNF_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct snf_handle snf_handle;

enum snf_status_code {
  SNF_OK = 0,
  SNF_BAD_PARAM = 1,     /* sizes or options the path does not support     */
  SNF_DEVICE_ERROR = 2,  /* a HIP call failed; the message says which       */
  SNF_NO_DEVICE = 3,     /* no gfx950 device visible                        */
  SNF_NOT_READY = 4      /* advance before upload                           */
};

enum snf_stop_code {     /* why an advance returned (cross.f90 istop/idump) */
  SNF_RUNNING = 0,       /* all requested steps done                         */
  SNF_STOP = 1,          /* istop: etime reached, NaN, laminar, dt < 1e-10   */
  SNF_DUMP = 2           /* idump: dumptime reached, caller should write     */
};

enum snf_stop_reason {   /* which istop condition ended the run (the first   */
  SNF_STOP_NONE = 0,     /* raised in it); the reference raises one flag     */
  SNF_STOP_ETIME = 1,    /* time reached etime (cross.f90:1699, 1709)        */
  SNF_STOP_DT = 2,       /* diverged: Deltat < 1e-10 (cross.f90:1682-1690)   */
  SNF_STOP_NAN = 3,      /* NaN in the UNER record (cross.f90:2102-2105)     */
  SNF_STOP_LAMINAR = 4   /* v' / Lz < 1e-5, laminarised (cross.f90:2106-2111)*/
};

/* Parameters of modules ctes/running/bcs (mod.f90) and hre2.dat/hre3.dat
 * (initcr.f90:54-135, 1950-1971) that the time advance reads.            */
typedef struct snf_params {
  int32_t mgalx, my, mgalz;   /* physical collocation points (hre2.dat)   */
  double Re;                  /* 1/nu                                     */
  double alp, gam;            /* fundamental wavenumbers, Lx=2pi/alp ...  */
  double Ly;                  /* box height                               */
  double s;                   /* mean shear rate                          */
  double chi;                 /* rotation (enters u x omega as 2 chi)     */
  double CFL;                 /* CFL number for the adaptive dt           */
  int32_t explicit_visc;      /* 1: explicit viscous terms (default);
                                 0: implicit Helmholtz stages             */
  int32_t nhist;              /* statistics every nhist steps             */
  int32_t readflag;           /* 0: generated IC (first dt uses umax>=1)  */
  double uprim;               /* IC velocity scale (isotropic first dt)   */
  double vbulk;               /* v of the (0,0) mode (cross.f90:1276)     */
  double etime, dumptime, dtimag;  /* stop / dump times (cross.f90:1680-1712) */
  int32_t precision;          /* SNF_PREC_FP64 (0), or SNF_PREC_MIXED (1):
                                 FP32 x-z transforms, hybrid planes and
                                 u x omega; FP64 state, KMM reconstruction,
                                 y operators, RK update, statistics sums;
                                 any grid (wave or Stockham x-z kernels) */
  int32_t device;             /* HIP device ordinal                       */
  int32_t ydealias;           /* iydealiasing (cross.f90:459-491), applied
                                 to vor, phi, U, W every substep: 0 none
                                 (default); 1 y-Fourier cut of the kx = 0
                                 lines and of U, W; 2 shifted cuts of every
                                 line in the moving frame, offsets wrapped
                                 to one period; 3 compact filter          */
} snf_params;

#define SNF_PREC_FP64 0
#define SNF_PREC_MIXED 1

/* Status returned by an advance: the scalars of module `running`/`bcs`. */
typedef struct snf_run_status {
  int32_t steps_done;
  int32_t stop;               /* snf_stop_code                             */
  double time, Deltat, xoffb, xofft;
  double umax[4];             /* umax(0:3) of the last CFL evaluation      */
  double CFLv;
  int32_t nhist_records;      /* UNER records produced by this call        */
  int32_t stop_reason;        /* snf_stop_reason when stop == SNF_STOP     */
} snf_run_status;

/* Number of doubles in one UNER record (.cfbin row, cross.f90:2087):
 * time, Deltat, uner(1:10), CFL, CFLv, umax(1:3), xoffb, xofft, umax(0),
 * uner2(1:4). */
#define SNF_UNER_LEN 24

/* ---- lifetime ---------------------------------------------------------
 * Replaces initcr(0) grid setup + derivadas/ffti (initcr.f90:186-306,
 * diffy_shear_hepta.f90:535-573, fou3D_thread.f:1-19).                  */
SNF_API int snf_create(const snf_params* p, snf_handle** out);
SNF_API int snf_destroy(snf_handle* h);
SNF_API const char* snf_last_error(void);

/* sizes: kxn = mx/2 complex kx modes, mz kz modes, nc = my*kxn*mz */
SNF_API int snf_sizes(const snf_handle* h, int32_t* kxn, int32_t* mz, int64_t* nc);

/* ---- state transfer (host jik layout, FP64) ---------------------------
 * Replaces the argument passing of cross (vor, phi, u00, w00) and the
 * module scalars time (running) and xoffb/xofft (bcs).                   */
SNF_API int snf_state_upload(snf_handle* h, const double* vor, const double* phi,
                     const double* u00, const double* w00,
                     double time, double xoffb, double xofft);
SNF_API int snf_state_download(snf_handle* h, double* vor, double* phi,
                       double* u00, double* w00,
                       double* time, double* xoffb, double* xofft);
/* same, in the device layout [y][kx][kz] (interleaved complex)           */
SNF_API int snf_state_upload_ykxkz(snf_handle* h, const double* vor, const double* phi,
                           const double* u00, const double* w00,
                           double time, double xoffb, double xofft);
SNF_API int snf_state_download_ykxkz(snf_handle* h, double* vor, double* phi,
                             double* u00, double* w00);

/* ---- the time advance -------------------------------------------------
 * Advances up to nstep RK3 steps (cross.f90:336-1029); returns early when
 * istop or idump is raised.  Each call is one `call cross` (its own step
 * count: nhist phase, the reference's unchecked second step), except a
 * call that follows an early return at a dump (stop == SNF_DUMP with fewer
 * than nstep steps done): it continues that run, as the reference's loop
 * continues after writing the dump.  snf_state_upload* starts a new run.
 * UNER records are kept on the device; fetch them with snf_history.      */
SNF_API int snf_advance(snf_handle* h, int32_t nstep, snf_run_status* st);
/* ---- shooting (the Newton/Arnoldi glue; the Krylov packing stays with
 * the caller) -------------------------------------------------------------
 * A state vector is vor, phi in the layout [y][kx][kz] (interleaved complex,
 * 2*my*kxn*mz doubles each) followed by u00, w00 (my doubles each).      */
SNF_API int snf_state_len(snf_handle* h, int64_t* n);
/* the shooting map of calc_dns_disk3d (gmres_shear.f90:2951-3120): from x
 * at time t0 with offsets xoffb0, xofft0, restart the statistics and run one
 * cross call to t0 + T (the etime stop); y receives the final state.    */
SNF_API int snf_shoot(snf_handle* h, const double* x, double t0, double xoffb0, double xofft0,
              double T, double* y, snf_run_status* st);
/* forward-difference Jacobian-vector products of the shooting map
 * (matvec_dns_disk3d with iopt_matvec=1, gmres_shear.f90:2666-2730):
 * out[d] = (Phi_T(x0 + eps[d] dirs[d]) - y0) / eps[d] for ndir directions
 * (state vectors, contiguous); y0 = Phi_T(x0), computed here if null.
 * The directions shoot concurrently: h and up to 15 copies of it (made on
 * first need while half the free device memory holds them, kept until
 * snf_destroy(h)), each bitwise the same as a separate snf_shoot.  h is left
 * holding the state of one of the shootings.                             */
SNF_API int snf_jvp_batch(snf_handle* h, const double* x0, const double* y0, const double* dirs,
                  int32_t ndir, const double* eps, double t0, double xoffb0, double xofft0,
                  double T, double* out);
/* ensembles: advances n independent handles by up to nstep steps each, as
 * n snf_advance calls would, with every handle's steps in flight on its own
 * stream before any is waited for (no counterpart in the reference, which
 * runs one box per MPI job).  st: n statuses (or null).                  */
SNF_API int snf_advance_batch(snf_handle* const* hs, int32_t n, int32_t nstep, snf_run_status* st);
/* copies up to max_records records (SNF_UNER_LEN doubles each) produced
 * since the last call; *n receives the count                             */
SNF_API int snf_history(snf_handle* h, double* out, int32_t max_records, int32_t* n);
/* per-plane statistics stats(1:23, 0:my-1) and stats2 (cross.f90:1798-2045),
 * column-major like the reference: out[s + 23*j]                          */
SNF_API int snf_statistics(snf_handle* h, double* stats, double* stats2,
                   int32_t* nacum, double* total_dt);
/* spectra of the statistics sp, sp2 (cross.f90:1970-2016): out[q + 8*(i +
 * kxn*k)] = sp(q+1, i, k), the sums over samples and planes of |u|^2,
 * |v|^2, |w|^2, Re(u v*), |o1|^2, |o2|^2, |o3|^2, Re(o1 o2*) with kx > 0
 * counted twice (sp2: each sample times its step)                        */
SNF_API int snf_spectra(snf_handle* h, double* sp, double* sp2);
/* zero stats, stats2, sp, sp2, nacum and total_dt, as escru does after
 * writing them (readwrite.f90:1322-1334)                                 */
SNF_API int snf_reset_statistics(snf_handle* h);

/* ---- operator entry points for parity tests (host arrays, [y][line]) --
 * derivyc (diffy_shear_hepta.f90:333-450): order 1 or 2 on nl lines of
 * length my; line l uses twist shp[l], shm[l].                           */
SNF_API int snf_op_derivy(snf_handle* h, const double* u, double* du, int32_t nl,
                  const double* shp, const double* shm, int32_t order);
/* lapcdy (diffy_shear_hepta.f90:466-525): (d_yy - k2[l]) p = u           */
SNF_API int snf_op_helmholtz(snf_handle* h, const double* u, double* p, int32_t nl,
                     const double* k2, const double* shp, const double* shm);
/* getv (operators.f90:735-793) on fields in the device layout [y][kx][kz]
 * (interleaved complex, my*kxn*mz values) with the twist of the handle's
 * current offsets: dir=+1 v = lapcdy(phi) (phi -> v), dir=-1 phi =
 * d2v/dy2 - k2 v (v -> phi).  The (0,0) line of out is not written.     */
SNF_API int snf_op_getv(snf_handle* h, const double* in, double* out, int32_t dir);
/* energy (operators.f90:870-924) of the device state: out[0..3] = sums of
 * |vor|^2, |phi|^2 (kx > 0 counted twice), U^2, W^2.                    */
SNF_API int snf_energy(snf_handle* h, double* out);
/* The same sums as snf_energy, from the same kernels (bitwise equal
 * results), with no device memory allocated or freed by the call: no
 * counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial sums live in a static device array of 16,416
 * bytes (2 x 1024 partials + 4 sums) that every GPU the library runs on
 * holds until the process ends; its size does not depend on the box, so no
 * box exceeds it.  Calls on one GPU share that array: the library
 * serialises them within a process, and every partial is written before it
 * is read, so a call depends on no earlier call's contents.            */
SNF_API int snf_energy_noalloc(snf_handle* h, double* out);
/* fourxz (fou3D_thread.f:22-73) on nplanes planes: dir=+1 spectral
 * (kxn x mz complex, [kx][kz]) -> physical (mgalx x mgalz, [x][z]);
 * dir=-1 the reverse with the 1/(mgalx mgalz) scaling.                   */
SNF_API int snf_op_fourxz(snf_handle* h, const double* in, double* out,
                  int32_t nplanes, int32_t dir);
/* one hvhg evaluation (cross.f90:1127-1765) on the uploaded state with the
 * given v, dv/dy and d(omega_y)/dy fields (device layout); returns the
 * three outputs (ome1c -> dvordy, rhgc -> hg, rhvc -> hv) and rf0u, rf0w. */
SNF_API int snf_op_hvhg(snf_handle* h, const double* hg_v, const double* hv_dv,
                const double* dvordy, double* out_dvordy, double* out_hg,
                double* out_hv, double* rf0u, double* rf0w);

/* ---- kernel timing (HIP events on the handle's stream) -----------------
 * No counterpart in the reference (its TIMING buckets, cross.f90:1041-1047,
 * are host wall-clock).  When enabled, snf_advance brackets every kernel
 * launch with events; snf_get_timing returns, per kernel class, the summed
 * milliseconds and the launch count since the last reset.               */
enum snf_kernel_class {
  SNF_K_XZ_A = 0,      /* spectral -> hybrid, inverse z DFT              */
  SNF_K_XZ_B = 1,      /* x c2r, u x omega, x r2c                        */
  SNF_K_XZ_C = 2,      /* hybrid -> spectral, forward z DFT, combination */
  SNF_K_Y = 3,         /* fused y-line substep (explicit or implicit)    */
  SNF_K_SMALL = 4,     /* scalar/CFL, twist tails, hist, step control    */
  SNF_K_NCLASS = 5
};
SNF_API int snf_set_timing(snf_handle* h, int32_t on);
SNF_API int snf_get_timing(snf_handle* h, double* ms, int32_t* counts, int32_t nclass);

/* Streaming-copy probe (no counterpart in the reference): copies nbytes
 * between two device buffers iters times with a plain grid-stride kernel
 * and returns the mean kernel time in *ms.  Used to measure attainable HBM
 * bandwidth and to calibrate the FETCH_SIZE / WRITE_SIZE counters.        */
SNF_API int snf_copy_probe(int32_t device, int64_t nbytes, int32_t iters, double* ms);

#ifdef __cplusplus
}
#endif
#endif
