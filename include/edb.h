/*
 * edb.h -- C ABI of the MI355X CalQ collision-operator library
 * (libedb.so, built from entropydecreaseboltzmann_amd/csrc).
 *
 * It replaces the reference's hot path, CalQ/QCalculator.h (genB, calQ),
 * CalQ/fft3d.h (fft, ifft), CalQ/EntropyFix.h (EntropyCal, setAverage,
 * FixSol) and the explicit drivers of CalQ/BKW.h (FEuler, RK2, SSP3, RK4,
 * CFE, DUFORT), with hand-written HIP kernels for gfx950.  The reference is
 * header-only C++ templates; each declaration below names the member it
 * stands in for (file:line in the reference repository).
 *
 * Conventions
 *   - Every function returns an int status (EDB_OK == 0); edb_last_error()
 *     gives a thread-local message for the last failure on the calling
 *     thread.  No C++ exception crosses this boundary.
 *   - Complex arrays are interleaved (re, im) doubles, layout-compatible with
 *     fftw_complex / std::complex<double> as the reference assumes
 *     (BKW.h:187-189).  Spectral arrays are N^3 entries in FFTW order.
 *   - Host-pointer entry points copy to and from the device and return when
 *     the result is on the host.  Device-pointer entry points take a
 *     hipStream_t passed as void* (NULL = the library's stream of the handle).
 *   - Handles own their device memory.  A handle must not be used from two
 *     threads at once; different handles may be used concurrently.
 *   - Grids: N in [3, 64] for the plain collision operator and the solver's
 *     plain precisions (stored or compressed beta), [3, 33] for the batched
 *     FP32 kernel
 *     (compiled for N in {3..13, 15, 16, 17, 21, 23, 24, 25, 31, 32, 33}),
 *     any 1 <= Nx, Ny, Nz <= 64 for the transforms.  For odd N the operator
 *     is the reference's; for even N (undefined in the reference) the
 *     Galerkin set is the 2n+1 = N-1 modes -n..n, n = (N-1)/2, differences
 *     wrap modulo N, and a term whose wrapped partner is the Nyquist index is
 *     absent (so Q_0 == 0 exactly and Q is Hermitian for Hermitian input).
 */
#ifndef EDB_H_
#define EDB_H_

#ifdef __cplusplus
extern "C" {
#endif

#define EDB_OK 0
#define EDB_EINVAL 1      /* bad argument (size, N, alpha, layout) */
#define EDB_ENOMEM 2      /* device memory exhausted */
#define EDB_EHIP 3        /* HIP runtime error */
#define EDB_ENOTREADY 4   /* genB not run / state not initialised */
#define EDB_EUNSUPPORTED 5
#define EDB_EBUSY 6

#define EDB_PREC_FP64_PLAIN 0    /* one instance at a time, FP64 beta streamed */
#define EDB_PREC_FP32_BATCHED 1  /* all instances in one launch, FP32 beta */
#define EDB_PREC_FP32_PLAIN 2    /* one instance at a time, FP32 beta streamed, FP64 products */
#define EDB_PREC_FP64_COMPRESSED 3 /* one instance at a time, FP64 beta formed on the fly from the
                                      compressed tables (edb_qcalc_genB_compressed) */

#define EDB_SCHEME_FEULER 0  /* BKW3D::FEuler, BKW.h:191-251 */
#define EDB_SCHEME_RK2 1     /* BKW3D::RK2,    BKW.h:314-384 */
#define EDB_SCHEME_SSP3 2    /* BKW3D::SSP3,   BKW.h:387-449 */
#define EDB_SCHEME_RK4 3     /* BKW3D::RK4,    BKW.h:452-519 */
#define EDB_SCHEME_CFE 4     /* BKW3D::CFE,    BKW.h:254-313 */
#define EDB_SCHEME_DUFORT 5  /* BKW3D::DUFORT, BKW.h:522-633 */

#define EDB_FIX_NONE 0        /* no entropy fix */
#define EDB_FIX_DIAGNOSTIC 1  /* the reference's: fixed F reported, state untouched */
#define EDB_FIX_APPLIED 2     /* fixed F transformed back into the spectral state */

/* ------------------------------------------------------------- library -- */

const char* edb_last_error(void);
int edb_version(int* major, int* minor);
/* Number of visible GPUs (0 without one); does not create a context. */
int edb_device_count(int* count);
/* Synchronise the device (all handles' streams). */
int edb_sync(void);

/* ------------------------------------------------- QCalculator (L2) ---- */

typedef struct edb_qcalc edb_qcalc;

/* QCalculator(int N, int alpha), QCalculator.h:39 + initialize() :43-52.
 * L is the velocity half-width that QCalculator.h:176 hard-codes as
 * 6.62132.  The Jackson filter starts on, as IS_FILTER = 1 at
 * QCalculator.h:34.  alpha: 0 Maxwell molecules, 1 hard spheres.  N in [3, 64]; the
 * batched kernels (genB_batched) and the solver exist up to N = 33, the plain
 * paths up to 64. */
int edb_qcalc_create(int N, int alpha, double L, edb_qcalc** out);

/* closeFilter(), QCalculator.h:54.  Must precede genB. */
int edb_qcalc_close_filter(edb_qcalc* q);

/* genB(), QCalculator.h:56, 158-229.  Builds beta on the device in FP64 in
 * the plain kernel's tiled layout (M^6 entries). */
int edb_qcalc_genB(edb_qcalc* q);

/* The FP32 paired layout used by the batched kernel (M^6/2 entries); no
 * reference counterpart (the reference has only the FP64 N^6 array). */
int edb_qcalc_genB_batched(edb_qcalc* q);
/* FP32 copy of genB's [l][k] beta (the FP64 values rounded once) for the
 * single-instance FP32-beta path: half the bytes of the plain HBM stream,
 * products and sums in FP64.  No reference counterpart (its B is FP64). */
int edb_qcalc_genB_plain32(edb_qcalc* q);
/* Compressed beta for the plain FP64 path, in place of genB's M^6 array:
 * beta(l,m) = F(lambda|l+m|, lambda|l-m|) sigma_l sigma_m rval - Bmm[m]
 * (QCalculator.h:199-228) depends on (l, m) only through |l+m|^2, |l-m|^2,
 * the filter factors and Bmm[m], so F is tabulated over those two integers
 * ((12n^2+1) x (6n^2+1) doubles: 8.5 MB at N=24, 0.53 GB at N=64) beside
 * M^3 (sigma, Bmm) pairs, and the kernel forms each beta on the fly.  Equal
 * to genB's beta bit for bit.  No reference counterpart (QCalculator.h
 * stores the N^6 array). */
int edb_qcalc_genB_compressed(edb_qcalc* q);
/* Device bytes of the compressed tables for q's N. */
int edb_qcalc_compressed_bytes(const edb_qcalc* q, long* bytes);

/* calQ(Qout, fin), QCalculator.h:59, 231-257: one Q-evaluation of an
 * N^3 complex spectrum in FFTW order, FP64 plain kernel.  Host arrays. */
int edb_qcalc_calQ(edb_qcalc* q, double* qout, const double* fin);
/* calQ with the FP32 copy of beta (edb_qcalc_genB_plain32). */
int edb_qcalc_calQ_plain32(edb_qcalc* q, double* qout, const double* fin);
/* calQ with beta formed from the compressed tables (edb_qcalc_genB_compressed),
 * with no M^6 array: every beta equals genB's, the sum runs in another
 * (fixed) order, so the result equals edb_qcalc_calQ to FP64 rounding.  Host
 * and device arrays as edb_qcalc_calQ / edb_qcalc_calQ_device. */
int edb_qcalc_calQ_compressed(edb_qcalc* q, double* qout, const double* fin);
int edb_qcalc_calQ_compressed_device(edb_qcalc* q, double* d_qout, const double* d_fin, void* stream);

/* Same on device arrays (FFTW-order N^3 complex). */
int edb_qcalc_calQ_device(edb_qcalc* q, double* d_qout, const double* d_fin, void* stream);

/* calQ over a batch sharing this beta with the FP32 batched kernel (needs
 * genB_batched).  Host arrays [batch][N^3] complex, FFTW order.  hermitian
 * != 0 evaluates half the k rows and mirrors them, exact only for spectra of
 * real data. */
int edb_qcalc_calQ_batched(edb_qcalc* q, int batch, int hermitian, double* qout, const double* fin);

/* Device variant: d_fin float2 [M^3 centred][ld], d_qout double2
 * [M^3 centred][ld], ld a multiple of 64 >= batch. */
int edb_qcalc_calQ_batched_device(edb_qcalc* q, int batch, int ld, int hermitian, double* d_qout,
                                  const float* d_fin, void* stream);

/* edb_qcalc_calQ_batched_device with no allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments and results (equal to
 * rounding of the FP64 merge order, bitwise when both use the same persistent
 * count); calls no hipMalloc or hipFree.  Its scratch (the work schedule's
 * prefix table and the stream-K partial slots) is a 64 MiB file-scope device
 * array, one per GPU, kept until the process ends; the persistent count is
 * capped to what it holds, and EDB_EINVAL is returned when even one wave's
 * slots and the prefix table (8 bytes per schedule item: at most ~9 KB per 64
 * instances, N = 33 with the rows schedule) do not fit.
 * Calls that overlap on one GPU (other streams, or a call still in flight on
 * another stream) share the array and race; calls in order on one stream do
 * not.  The first call on a device looks up the array's address there. */
int edb_qcalc_calQ_batched_device_noalloc(edb_qcalc* q, int batch, int ld, int hermitian, double* d_qout,
                                          const float* d_fin, void* stream);

/* beta in the reference's layout B[l][k] (M^6 doubles, l and k centred,
 * QCalculator.h:199-224), computed on the device; for verification. */
int edb_qcalc_copy_B(edb_qcalc* q, double* Bout);

/* N, modes per dimension M = 2n+1, and bytes of each resident beta layout. */
int edb_qcalc_info(const edb_qcalc* q, int* N, int* M, long* bytes_plain64, long* bytes_pair32);

/* Time the batched kernel alone: reps launches on the handle's stream,
 * bracketed by HIP events, after ~40 ms of warm-up launches (the device's
 * clock settles over the first launches after idle); returns the mean
 * milliseconds per launch. */
int edb_qcalc_time_batched(edb_qcalc* q, int batch, int hermitian, int reps, double* ms);

/* Same for the plain FP64 kernel (one instance). */
int edb_qcalc_time_plain(edb_qcalc* q, int reps, double* ms);
/* Same for the plain kernel with beta formed from the compressed tables. */
int edb_qcalc_time_compressed(edb_qcalc* q, int reps, double* ms);

/* Launch-shape knobs: persistent waves of the batched kernel and l-slabs of
 * the plain kernel (0 = built-in default).  No reference counterpart. */
int edb_qcalc_tune(edb_qcalc* q, int batched_waves, int plain_slabs);

/* Autotuner: times the candidate launch shapes of the kernels built so far
 * (batched at `batch` instances, Hermitian half) with HIP events and keeps
 * the fastest; returns them.  No reference counterpart. */
int edb_qcalc_autotune(edb_qcalc* q, int batch, int* batched_waves, int* plain_slabs);

/* Plain-kernel knobs of the stored-beta stream, FP64 beta (fp32 == 0) or its
 * FP32 copy (fp32 != 0): the kernel variant (0: f_[k-l] gathered from the
 * difference table by one vector load per term, beta in groups of 8; 1, 2,
 * 3: f_[k-l] from a per-l-row window of the table staged in LDS, beta in
 * groups of 4, 2, 8; -1 keeps it) and
 * the l-slab count (0 keeps it).  Every variant sums in the same order with
 * the same arithmetic, so results are bitwise equal.  The autotuner times
 * them.  No reference counterpart. */
int edb_qcalc_tune_plain(edb_qcalc* q, int fp32, int kernel, int slabs);
int edb_qcalc_plain_schedule(const edb_qcalc* q, int fp32, int* kernel, int* slabs);

/* Work schedule of the batched kernel.  EDB_SCHED_ROWS: stream-K over
 * (k-row, instance group, l-row) steps with `count` persistent waves.
 * EDB_SCHED_QUADS: workgroups of 4 neighbouring k-rows walking the union of
 * their l-row schedules together, so the f rows they share come from L2;
 * `count` persistent workgroups.  count 0 = built-in default.  The
 * autotuner times both and keeps the faster.  No reference counterpart. */
#define EDB_SCHED_ROWS 0
#define EDB_SCHED_QUADS 1
int edb_qcalc_tune_batched(edb_qcalc* q, int schedule, int count);
int edb_qcalc_batched_schedule(const edb_qcalc* q, int* schedule, int* count);

/* Conservative moment correction of every Q this handle returns (calQ,
 * calQ_device, calQ_batched, calQ_batched_device): the least-squares nearest
 * spectrum on the Galerkin modes whose discrete mass, momentum and energy
 * moments (h^3 sum_x {1, v, |v|^2} Q(x) on the grid) are zero.  Off by
 * default.  No counterpart in the reference, whose Q conserves mass only;
 * it is the constrained-L2 projection of conservative spectral methods. */
int edb_qcalc_set_conservative(edb_qcalc* q, int on);

int edb_qcalc_destroy(edb_qcalc* q);

/* ------------------------------------------------------ fft3d (L2) ----- */

/* fft3d::fft(Nx,Ny,Nz, re, fre), fft3d.h:25-30: forward, unnormalised,
 * exp(-2 pi i jk/N).  Host arrays of Nx*Ny*Nz complex. */
int edb_fft3d_fft(int Nx, int Ny, int Nz, const double* re, double* fre);

/* fft3d::ifft(Nx,Ny,Nz, re, fre), fft3d.h:32-48: backward then times
 * 1/(Nx Ny Nz).  Note the reference's argument order (output first). */
int edb_fft3d_ifft(int Nx, int Ny, int Nz, double* re, const double* fre);

/* Batched device transforms on instance-minor arrays [Nx*Ny*Nz][ld]
 * complex (ld a multiple of 64 >= batch); inverse != 0 applies the 1/N^3
 * scale.  In-place allowed for lengths <= 33. */
int edb_fft3d_batched_device(int Nx, int Ny, int Nz, int batch, int ld, int inverse,
                             const double* d_in, double* d_out, void* stream);

/* ------------------------------------------------- EntropyFix (L2) ----- */

/* The state of EntropyFix<i_type, r_type> (EntropyFix.h:16-22): size, eps,
 * the cached average of f and whether it is set. */
typedef struct {
  long Nt;
  double eps;
  double average;
  int average_flag;
} edb_efix_state;

/* EntropyFix(Nt, eps = 1e-10), EntropyFix.h:29-32. */
int edb_efix_init(edb_efix_state* st, long Nt, double eps);
/* setAverage(Nf, fin), EntropyFix.h:65-78. */
int edb_efix_set_average(edb_efix_state* st, long Nf, const double* f);
/* EntropyCal(Nf, fin), EntropyFix.h:124-161 (sets the average on first use). */
int edb_efix_entropy_cal(edb_efix_state* st, long Nf, const double* f, double* H);
/* FixSol(Nf, fin, Aim), EntropyFix.h:211-253: bisection toward the average,
 * f overwritten.  Runs on the GPU. */
int edb_efix_fixsol(edb_efix_state* st, long Nf, double* f, double aim);

/* ------------------------------------------------------ BKW3D (L3) ----- */

typedef struct {
  int N;           /* grid points per dimension */
  int alpha;       /* 0 Maxwell molecules, 1 hard spheres */
  double L;        /* velocity half-width, domain [-L, L]^3 */
  int filter;      /* Jackson filter on (reference default) or off */
  int batch;       /* instances */
  int precision;   /* EDB_PREC_* */
  int scheme;      /* EDB_SCHEME_* */
  double dt;
  double beta;     /* CFE / DUFORT relaxation parameter (BKW.h:255, 523) */
  int fix_mode;    /* EDB_FIX_* */
  int hermitian;   /* batched kernel evaluates half the k rows (real data) */
  int history;     /* steps of entropy history kept on the device (0: none) */
  int conserve;    /* moment correction of every Q (edb_qcalc_set_conservative) */
} edb_solver_params;

typedef struct edb_solver edb_solver;

/* BKW3D(N0, R), BKW.h:39 + the QCalculator/EntropyFix members each driver
 * creates (BKW.h:462-473).  A qcalc built for the same (N, alpha, L,
 * filter) may be shared (beta reused); pass NULL to build one. */
int edb_solver_create(const edb_solver_params* p, edb_qcalc* shared, edb_solver** out);

/* setInitialValue(T0) + fft3(), BKW.h:121-138, 182-190, with caller data:
 * F real, [batch][N^3], grid x_i = -L + i*2L/N. */
int edb_solver_set_initial_real(edb_solver* s, double T0, const double* F);
/* Spectral initial data [batch][N^3] complex, FFTW order. */
int edb_solver_set_initial_spectral(edb_solver* s, double T0, const double* Fhat);

/* Advance nsteps steps of the scheme (one pass of the driver's do-loop body
 * each: stages, ifft, calEntropy, entropy fix). */
int edb_solver_step(edb_solver* s, int nsteps);
/* The drivers' loop do { t += dt; ... } while (t < TEnd), BKW.h:478-516. */
int edb_solver_run(edb_solver* s, double TEnd, int* nsteps);

/* Record one step into a HIP graph and replay it for later steps (after the
 * first step, whose entropy-fix eligibility differs). */
int edb_solver_use_graph(edb_solver* s, int enable);

/* Mean milliseconds per Q-kernel launch over the steps since the last call
 * (events around each launch; only counted when graphs are off), and the
 * number of launches. */
int edb_solver_qkernel_time(edb_solver* s, double* ms_mean, long* launches);
int edb_solver_enable_qtiming(edb_solver* s, int enable);

int edb_solver_time(const edb_solver* s, double* t, int* steps);
/* [batch][N^3] complex spectral state, FFTW order. */
int edb_solver_get_spectral(edb_solver* s, double* Fhat);
/* [batch][N^3] real part of ifft of the state (the driver's F). */
int edb_solver_get_real(edb_solver* s, double* F);
/* Entropy history: H after each step, after its fix, and whether a fix ran,
 * [steps][batch] for the first min(steps, history) steps. */
int edb_solver_entropy_history(edb_solver* s, int steps, double* H, double* Hfix, int* fixed);
/* Mean of Re F after each step, before and after its entropy fix (the
 * drivers' "average" column, BKW.h:288-305, 357-378), [steps][batch]. */
int edb_solver_mean_history(edb_solver* s, int steps, double* mean, double* mean_fix);
/* The driver's real-space F after the last step (BKW3D::F after fft3(true)
 * and the fix, BKW.h:485-512): the entropy fix's values where it ran, in the
 * diagnostic mode too, whose fixed F never reaches the state.  [batch][N^3]. */
int edb_solver_get_driver_real(edb_solver* s, double* F);
/* FP64 moments of F: [batch][5] = mass, momentum x,y,z, energy (sum |v|^2 f). */
int edb_solver_moments(edb_solver* s, double* out);
/* L-infinity and L1 error of F against the BKW profile at K, [batch][2]. */
int edb_solver_bkw_error(edb_solver* s, double K, double* out);

int edb_solver_destroy(edb_solver* s);

/* ------------------------------------------- request server (serving) -- */
/* Continuous batching of independent solves (RK4, dt, entropy fix NONE or
 * APPLIED).  Requests of one (N, alpha) share a group: a cached collision
 * kernel and `max_slots` device-resident instance slots on a stream of their
 * own; finished slots are freed and waiting requests admitted at chunk
 * boundaries.  No reference counterpart (the reference has no serving). */
typedef struct {
  double L;          /* velocity half-width */
  int filter;        /* Jackson filter */
  double dt;
  int fix_mode;      /* EDB_FIX_NONE or EDB_FIX_APPLIED */
  int max_slots;     /* slots per group (0: 256, 64 for N >= 27) */
  int max_chunk;     /* RK4 steps per launch round (0: 16) */
  long hbm_budget;   /* bytes for kernels + slots over all groups (0: unlimited) */
  int conserve;      /* moment correction of every Q (edb_qcalc_set_conservative) */
  int max_pending;   /* admission limit: submit returns EDB_EBUSY while this many
                        requests wait for a slot (0: unlimited) */
} edb_server_params;

typedef struct edb_server edb_server;

int edb_server_create(const edb_server_params* p, edb_server** out);
/* Queue one solve of f(t=0) = F0 (N^3 real on x_i = -L + i*2L/N) to t_end
 * (the drivers' step count, do { t += dt } while (t < t_end)).  EDB_ENOMEM:
 * its group can never fit the budget; EDB_EBUSY: not now (groups in use);
 * EDB_EINVAL: bad arguments, including a non-finite value in F0 and
 * t_end / dt above 1e8 steps. */
int edb_server_submit(edb_server* s, long id, int N, int alpha, double t_end, const double* F0);
/* Advance the scheduler without blocking: retire finished chunks, admit
 * waiting requests, launch the next chunk of each group.  *ready = number of
 * results waiting to be fetched. */
int edb_server_poll(edb_server* s, int* ready);
/* Pop one finished solve: its id, N, steps, final t (the t accumulated by the
 * drivers' loop, which may pass t_end by up to dt, as edb_solver_time reports)
 * and real-space F. */
int edb_server_fetch(edb_server* s, long* id, int* N, int* steps, double* t, double* F, long capacity);
int edb_server_stats(const edb_server* s, long* submitted, long* finished, long* active, long* waiting,
                     long* hbm_used);
int edb_server_destroy(edb_server* s);

#ifdef __cplusplus
}
#endif
#endif
