/*
 * mafx.h - C-ABI of the MI355X-native magnum.af effective-field / integrator hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an int status
 * (MAFX_OK == 0).  On failure mafx_last_error() returns a thread-local message.  No C++
 * exception crosses this boundary.
 *
 * Layout of every vector field (m, H, k, gradients): structure of arrays, double or float,
 *     index(c, x, y, z) = x + nx * (y + ny * (z + nz * c)),   c in {0,1,2}
 * which is bit-for-bit the column-major af::array [nx, ny, nz, 3] the reference passes around
 * (magnumaf/include/magnumaf/mesh.hpp:34, "dims_v").  Scalar fields (Ms_field, rkky values)
 * use index(x, y, z) = x + nx * (y + ny * z).
 *
 * Pointers named *_dev are device (HBM) pointers; pointers named *_host are host pointers
 * that are read during the call only.  A context owns all of its device scratch memory and
 * launches on one HIP stream (the null stream unless mafx_ctx_set_stream is called).
 * A context is externally synchronised: one host thread at a time.
 */
#ifndef MAFX_H
#define MAFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAFX_OK 0
#define MAFX_ERR_ARG 1      /* invalid argument (sizes, null pointers, unsupported combination) */
#define MAFX_ERR_DEVICE 2   /* a HIP runtime call failed */
#define MAFX_ERR_NUMERIC 3  /* non-finite value in a reduction / line search failure */
#define MAFX_ERR_STATE 4    /* object used in the wrong state */

/* Mesh: magnumaf/include/magnumaf/mesh.hpp:11-20 (nx, ny, nz cells of dx x dy x dz metres). */
typedef struct mafx_mesh {
    int64_t nx, ny, nz;
    double dx, dy, dz;
} mafx_mesh;

/* Arithmetic of the field evaluation.  MAFX_F64 is the plain FP64 path (the reference's
 * arithmetic).  MAFX_F32_FIELD runs the demag FFT, tensor product and stencils in FP32 with
 * FP32 accumulation; the state m, the RK combinations, error norm, renormalisation and all
 * energy / dot-product sums stay FP64.  MAFX_F32_DEMAG runs only the demag (or dipole) FFT
 * convolution in FP32 and the local terms in FP64; MAFX_F32_STENCILS the reverse (FP64 demag,
 * FP32 local terms), which exists to separate the two sources of FP32 error. */
enum mafx_precision { MAFX_F64 = 0, MAFX_F32_FIELD = 1, MAFX_F32_DEMAG = 2, MAFX_F32_STENCILS = 3 };

/* Field terms.  Each kind replaces one reference FieldTerm subclass; the cited file holds
 * the impl_H_in_Apm / impl_E_in_J the GPU evaluation reproduces. */
enum mafx_term_kind {
    MAFX_DEMAG = 1,          /* DemagField(mesh)            field_terms/micro/demag_field.cpp:211-264 */
    MAFX_EXCHANGE = 2,       /* ExchangeField(A)            field_terms/micro/exchange_field.cpp:14-60,115-157 */
    MAFX_DMI = 3,            /* DmiField(D, axis)           field_terms/micro/dmi_field.cpp:41-110 */
    MAFX_UNIAXIAL = 4,       /* UniaxialAnisotropyField     field_terms/micro/uniaxial_anisotropy_field.cpp:58-89 */
    MAFX_ZEEMAN = 5,         /* ExternalField(H)            field_terms/micro/external_field.cpp:28-48 */
    MAFX_RKKY = 6,           /* RKKYExchangeField           field_terms/micro/rkky_exchange_field.cpp:19-29,195-337 */
    MAFX_ATOM_DIPOLE = 11,   /* AtomisticDipoleDipoleField  field_terms/atom/atomistic_dipole_dipole_field.cpp:14-113 */
    MAFX_ATOM_EXCHANGE = 12, /* AtomisticExchangeField(J)   field_terms/atom/atomistic_exchange_field.cpp:8-40 */
    MAFX_ATOM_DMI = 13,      /* AtomisticDmiField(D, axis)  field_terms/atom/atomistic_dmi_field.cpp:18-49 */
    MAFX_ATOM_UNIAXIAL = 14, /* AtomisticUniaxialAnisotropyField  field_terms/atom/atomistic_uniaxial_anisotropy_field.cpp:24-34 */
    MAFX_ATOM_ZEEMAN = 15    /* AtomisticExternalField(H)   field_terms/atom/atomistic_external_field.cpp:5-11 */
};

/* Time-dependent field of an ExternalField callback (external_field.cpp:13-17,28-36, whose
 * callbacks take the State and return the field).  Called on the calling host thread at every
 * evaluation that includes the term, with the evaluation time t (the context time of
 * mafx_ctx_set_time, or t + c_i h at stage i of an integrator attempt) and the magnetisation being
 * evaluated (m, device, [3N] f64).  It writes the field in A/m into h_out and returns 0, any other
 * value fails the evaluation with MAFX_ERR_ARG.  h_out is the library's device buffer [3N] f64,
 * to be filled on `stream` (a hipStream_t) before the callback returns.  With
 * MAFX_TERM_HOMOGENEOUS_CALLBACK h_out is a host double[3] holding one homogeneous field; with
 * MAFX_TERM_HOST_CALLBACK m and h_out are host buffers [3N] (the library copies m down before the
 * call and the field up after it), for callers that compute the field on the host. */
typedef int (*mafx_field_fn)(void* user, double t, const double* m, double* h_out, void* stream);

/* One field term.  Unused members are ignored.
 *   value:   A [J/m] (EXCHANGE), D [J/m^2] (DMI), Ku1 [J/m^3] (UNIAXIAL),
 *            J_atom / D_atom / K_atom [J] (ATOM_*)
 *   axis:    DMI interface normal or anisotropy axis; normalised by the library as the
 *            reference constructors do (dmi_field.cpp:43-47, uniaxial_anisotropy_field.cpp:11-12)
 *   field:   homogeneous external field [A/m] (ZEEMAN / ATOM_ZEEMAN) when array0 is NULL
 *   array0:  ZEEMAN / ATOM_ZEEMAN: host vector field [3*N] in A/m (overrides `field`);
 *            RKKY: host scalar field of RKKY constants [N] (J/m, the reference's rkky_values);
 *            DMI: host scalar field of D [N] in J/m^2, overrides `value`
 *            (DmiField(af::array D_constants, axis), dmi_field.cpp:17-19,97-109);
 *            UNIAXIAL: host scalar field of Ku1 [N] in J/m^3, overrides `value`
 *            (UniaxialAnisotropyField(af::array Ku1_field, ...), uniaxial_anisotropy_field.cpp:14-16,78-88)
 *   array1:  RKKY: host scalar field of exchange constants A [N] (J/m);
 *            UNIAXIAL: host vector field [3*N] of anisotropy axes, overrides `axis`; each cell's
 *            axis is divided by its length as util::normalize does (uniaxial_anisotropy_field.cpp:23-25,60-69)
 *   index0:  RKKY: host scalar field of uint32 layer indices [N], or NULL for all zero
 *   callback, callback_user: ZEEMAN / ATOM_ZEEMAN: a time-dependent field (mafx_field_fn),
 *            overrides `field` and array0
 *   flags:   DEMAG / ATOM_DIPOLE: MAFX_TERM_CACHING reads the transformed tensor from the on-disk
 *            cache and writes it there when absent, MAFX_TERM_VERBOSE reports cache activity on
 *            stdout (the caching and verbose arguments of DemagField(mesh, verbose, caching),
 *            demag_field.cpp:185-204, cache_manager.cpp:49-114).  MAFX_TERM_HOST_TENSOR (no
 *            counterpart in the reference, whose only tensor is this one) and its companions choose
 *            the real-space tensor the transform starts from.  By default the FP64 paths use the
 *            reference's tensor over the whole expanded grid, evaluated on the host in the
 *            reference's operation order (bit for bit, far-field rounding noise included), and the
 *            FP32 paths the same evaluation on the non-negative octant extended by the components'
 *            parities (an eighth of the host work; it differs from the whole tensor only by the
 *            summation-order noise of the negative offsets).  MAFX_TERM_HOST_TENSOR selects the
 *            whole host tensor, MAFX_TERM_OCTANT_TENSOR the extended octant, and
 *            MAFX_TERM_DEVICE_TENSOR the same formulas evaluated on the GPU (milliseconds instead of
 *            seconds, with the device libm's own far-field noise), on either path.
 *            ZEEMAN / ATOM_ZEEMAN: MAFX_TERM_HOMOGENEOUS_CALLBACK, the callback writes one
 *            homogeneous field into a host double[3]; MAFX_TERM_HOST_CALLBACK, the callback
 *            reads m from and writes the field into host buffers
 */
#define MAFX_TERM_CACHING 1
#define MAFX_TERM_VERBOSE 2
#define MAFX_TERM_HOST_TENSOR 4
#define MAFX_TERM_HOMOGENEOUS_CALLBACK 8
#define MAFX_TERM_HOST_CALLBACK 16
#define MAFX_TERM_DEVICE_TENSOR 32
#define MAFX_TERM_OCTANT_TENSOR 64
/* DEMAG on the FP32 paths: apply the spectrum of the reference's whole tensor (far-field rounding
 * noise and its asymmetry included) instead of the parity-symmetric one.  It is stored as the
 * symmetric real quarter (FP32) plus the residual over the whole half spectrum (complex FP16
 * scaled per component by a power of two; the residual is ~1e-4 of the entries, so its rounding
 * leaves ~1e-7): 0.13 + 0.40 GB at 1024x1024x4 instead of 0.13 GB.  With
 * MAFX_TERM_DEVICE_TENSOR the GPU-built whole tensor takes the host tensor's place.  Ignored by
 * the FP64 paths. */
#define MAFX_TERM_FULL_TENSOR 128
typedef struct mafx_term_desc {
    int32_t kind;
    int32_t flags;
    double value;
    double axis[3];
    double field[3];
    const double* array0;
    const double* array1;
    const uint32_t* index0;
    mafx_field_fn callback;
    void* callback_user;
} mafx_term_desc;

typedef struct mafx_ctx mafx_ctx;
typedef struct mafx_llg mafx_llg;

/* ------------------------------------------------------------------ errors / device info */
const char* mafx_last_error(void);
/* Number of visible devices (0 without a GPU).  Does not create a context. */
int mafx_device_count(int* count);

/* ------------------------------------------------------------------ context */
/* Creates the evaluation context for one mesh, one material (Ms scalar, or Ms_field when
 * Ms_field_host != NULL, State ctor semantics of state.cpp:71-111), one term list and one
 * precision.  Uploads parameters, builds the demag / dipole tensor on the GPU in FP64
 * (demag_field.cpp:104-147), plans the FFTs and allocates the workspace.  Replaces the
 * construction of the reference's terms plus State::Ms/Ms_field.  Reads MAFX_NUM_CUS (CU count
 * for the persistent transform grids; default: the device's) and, with MAFX_TERM_CACHING,
 * MAFX_CACHE_DIR / MAFX_CACHE_MAX_BYTES. */
int mafx_ctx_create(const mafx_mesh* mesh, double Ms, const double* Ms_field_host,
                    const mafx_term_desc* terms, int nterms, int precision, mafx_ctx** out);
int mafx_ctx_destroy(mafx_ctx* ctx);
/* Launch on an existing HIP stream (hipStream_t passed as void*); NULL selects the null stream. */
int mafx_ctx_set_stream(mafx_ctx* ctx, void* stream);
int mafx_ctx_sync(mafx_ctx* ctx);
/* Device bytes held by the context (tensor, FFT workspace, parameters). */
int mafx_ctx_device_bytes(const mafx_ctx* ctx, int64_t* bytes);
/* Time of the State the next field and energy evaluations see (State::t, state.hpp), used by
 * time-dependent ExternalField callbacks only.  The integrator sets it per RK stage itself. */
int mafx_ctx_set_time(mafx_ctx* ctx, double t);
/* Replaces the parameters of term `index` (same kind, same mesh) in place: the values, axes and
 * arrays of desc are uploaded again, as assigning a term's members does in the reference
 * (UniaxialAnisotropyField::Ku1, ExternalField::set_homogeneous_field, external_field.cpp:21-26).
 * A DEMAG / ATOM_DIPOLE term keeps its tensor.  Integrators and minimizers of the context see the
 * new parameters from their next evaluation on. */
int mafx_ctx_update_term(mafx_ctx* ctx, int index, const mafx_term_desc* desc);
/* Replaces the material: Ms (A/m, or mu_s in J/T for atomistic terms) or, when Ms_field_host is
 * not NULL, the cell-wise Ms_field [N] (State::Ms, State::set_Ms_field, state.cpp:135,
 * magnumaf.pyx:777-784,839-842). */
int mafx_ctx_set_material(mafx_ctx* ctx, double Ms, const double* Ms_field_host);

/* Real-space Newell demag tensor on the expanded grid, [6][n2][n1][n0] with n0 = 2 nx,
 * n1 = 2 ny, n2 = (nz == 1 ? 1 : 2 nz), components (xx, xy, xz, yy, yz, zz), evaluated on the host
 * in the reference's operation order.  Replaces DemagField::calc_N (demag_field.cpp:104-142,207-209).
 * nthreads <= 0 uses every hardware thread.  Needs no GPU. */
int mafx_demag_calc_N(const mafx_mesh* mesh, double* N_host, int nthreads);

/* On-disk cache of transformed demag / dipole tensors (DemagField caching, demag_field.cpp:185-204;
 * CacheManager, cache_manager.cpp:49-114).  Folder: $MAFX_CACHE_DIR, else
 * <temp_directory_path>/magnumaf.cache as in the reference.  File name: the reference's mesh string
 * to_string(mesh) (demag_field.cpp:185-190) followed by the layout (precision, dipole) and the exact
 * bits of dx, dy, dz.  When the folder exceeds $MAFX_CACHE_MAX_BYTES (default 2e9, the reference's
 * max_size_in_byte) the oldest files are removed until it is below a tenth of that (2e8 by
 * default, the reference's shrink_size_in_byte).  A file that fails its header or checksum check
 * is ignored and the tensor computed instead, as the reference does when af::readArray fails.
 * mafx_demag_cache_name writes the file name for a mesh, precision and DEMAG term flags (the
 * MAFX_TERM_HOST_TENSOR / MAFX_TERM_DEVICE_TENSOR layouts have names of their own; needs no GPU); mafx_ctx_tensor_source
 * tells where a context's tensor came from: 0 no tensor, 1 computed, 2 read from the cache. */
int mafx_demag_cache_name(const mafx_mesh* mesh, int precision, int term_flags, int dipole, char* buf,
                          int64_t buf_len);
int mafx_ctx_tensor_source(const mafx_ctx* ctx, int* source);

/* ------------------------------------------------------------------ field and energy */
/* H_eff = sum of the selected terms, in term order (field_term.hpp:116-123), A/m.
 * term_mask bit i selects terms[i]; ~0u selects all.  h_dev is double* for MAFX_F64 and
 * double* or float* for MAFX_F32_FIELD according to h_is_f32.  Replaces
 * fieldterm::Heff_in_Apm and FieldTerm::H_in_Apm. */
int mafx_heff(mafx_ctx* ctx, const double* m_dev, uint32_t term_mask, void* h_dev, int h_is_f32);
/* Energy of every selected term (J) and their sum, each with the reference's formula:
 * micro -mu0/2 Ms dV sum(h.m) (micro_term.hpp:18-30), exchange with the Neumann-corrected
 * field (exchange_field.cpp:14-60), Zeeman without the 1/2 (external_field.cpp:38-48),
 * atomistic -mu0/2 mu_s sum(h.m) (atom_term.hpp:12-15, atomistic_external_field.cpp:5-11).
 * E_terms_host has nterms entries (0 for unselected terms).  Replaces
 * FieldTerm::Energy_in_J and fieldterm::Eeff_in_J. */
int mafx_energy(mafx_ctx* ctx, const double* m_dev, uint32_t term_mask, double* E_terms_host,
                double* E_total_host);

/* ------------------------------------------------------------------ vector helpers (FP64) */
/* k = dt * LLG(alpha, m, h): -gamma/(1+alpha^2) [m x h + alpha m x (m x h)]
 * (equations.hpp:19-33, llg_integrator.cpp:18-38); h is FP64. */
int mafx_llg_rhs(mafx_ctx* ctx, double alpha, double dt, const double* m_dev, const double* h_dev,
                 double* k_dev);
/* sum(a*b) over 3N entries, FP64 fixed-order tree (util.cpp:30-32 full_inner_product). */
int mafx_dot(mafx_ctx* ctx, const double* a_dev, const double* b_dev, double* out_host);
/* max |a| over 3N entries (math.cpp:6-8 max_4d_abs). */
int mafx_max_abs(mafx_ctx* ctx, const double* a_dev, double* out_host);
/* In-place normalisation (util.cpp:34 normalize_even_zero_vectors when zero_safe == 0,
 * util.cpp:37-42 normalize_handle_zero_vectors when zero_safe != 0). */
int mafx_normalize(mafx_ctx* ctx, double* m_dev, int zero_safe);
/* <m> over cells with |m| != 0 (state.cpp:32-45). */
int mafx_mean_m(mafx_ctx* ctx, const double* m_dev, double out_host[3]);
/* No counterpart in the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight output loops, harnesses that forbid hipMalloc): mafx_mean_m's result, bit for
 * bit, without a context.  m_dev is [3][ncells] f64 (SoA, as everywhere); out_dev receives 3
 * doubles in device memory, in stream order on `stream` (NULL: the null stream); nothing is
 * synchronised.  The reduction's partial sums live in a static 32 KiB device array (4 x 1024
 * doubles, one per GPU for the life of the process), so the call allocates nothing.  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream, or events).
 * ncells must be in [1, 2^60]; the partial array is the same size for every ncells. */
int mafx_mean_m_noalloc(const double* m_dev, int64_t ncells, double* out_dev, void* stream);

/* ------------------------------------------------------------------ adaptive RK / LLG */
/* LLGIntegrator(alpha, terms, scheme, Controller(hmin, hmax, atol, rtol), dissipation_term_only)
 * (llg_integrator.hpp:26-49, controller.hpp:11-17).  scheme: "RKF45", "DP45", "BS45", "DP78" or
 * "BS23" (adaptive_runge_kutta.cpp:8-37,75-529).  The integrator evaluates all terms of ctx. */
int mafx_llg_create(mafx_ctx* ctx, double alpha, const char* scheme, double hmin, double hmax,
                    double atol, double rtol, int dissipation_term_only, mafx_llg** out);
int mafx_llg_destroy(mafx_llg* llg);
int mafx_llg_set_alpha(mafx_llg* llg, double alpha);
/* Cell-wise damping (LLGIntegrator(af::array alpha_field, ...), llg_integrator.hpp:32-40,
 * llg_integrator.cpp:18-38): alpha_host [N] is copied to the device and replaces the scalar alpha
 * in every later right-hand side; NULL returns to the scalar alpha. */
int mafx_llg_set_alpha_field(mafx_llg* llg, const double* alpha_host);
/* One accepted step of AdaptiveRungeKutta::step (adaptive_runge_kutta.cpp:39-71): repeats
 * attempts until the controller accepts, then t += h, m += dm, renormalise.  m_dev is
 * updated in place, *t_inout advanced. */
int mafx_llg_step(mafx_llg* llg, double* m_dev, double* t_inout);
/* LLGIntegrator::relax (llg_integrator.cpp:43-66): while the relative energy change
 * |(E_prev - E) / E_prev| between two checks exceeds precision (E_prev starts at 1e20), takes
 * eval_E steps; with verbose, prints the relative change and <m> whenever the State's step count
 * (state_steps before the call plus the steps taken) is a multiple of iwritecout (> 0).  E is the
 * total energy of all terms of ctx. */
int mafx_llg_relax(mafx_llg* llg, double* m_dev, double* t_inout, double precision, int64_t eval_E,
                   int64_t iwritecout, int verbose, int64_t state_steps, int64_t* nsteps_done);
/* Runs mafx_llg_step until *t_inout >= t_end or nsteps_max steps were taken. */
int mafx_llg_run_until(mafx_llg* llg, double* m_dev, double* t_inout, double t_end,
                       int64_t nsteps_max, int64_t* nsteps_done);
/* LLGIntegrator::integrate_dense (llg_integrator.cpp:68-104): integrates time_in_s and writes
 * (t, <mx>, <my>, <mz>) rows at every write_every_dt, the first row at the start time; the
 * output sample comes from a backwards RKF45 step from the overshooting state.
 * out_host holds max_rows*4 doubles; *rows_written is set. */
int mafx_llg_integrate_dense(mafx_llg* llg, double* m_dev, double* t_inout, double time_in_s,
                             double write_every_dt, double* out_host, int64_t max_rows,
                             int64_t* rows_written);

typedef struct mafx_llg_stats {
    double h;                 /* next step size (AdaptiveRungeKutta::h_) */
    double h_last;            /* step size of the last accepted step */
    double err_last;          /* error norm of the last accepted attempt */
    int64_t accepted, rejected, heff_evals;
    int64_t counter_hmin, counter_hmax, counter_maxscale, counter_minscale;
} mafx_llg_stats;
int mafx_llg_get_stats(const mafx_llg* llg, mafx_llg_stats* out);
/* The last accepted attempt's error estimate sequence lives on the host; this restores
 * controller state for an exact resume (h, errold, reject flag). */
int mafx_llg_set_controller_state(mafx_llg* llg, double h, double errold, int reject);
int mafx_llg_get_controller_state(const mafx_llg* llg, double* h, double* errold, int* reject);

/* ------------------------------------------------------------------ LBFGS */
/* LBFGS_Minimizer(terms, tolerance, maxIter, verbose)::Minimize (lbfgs_minimizer.cpp:55-198):
 * L-BFGS (history 5) with the More-Thuente line search on the sphere (cvsrch/cstep,
 * :213-505).  m_dev is relaxed in place; returns the final energy in *E_out.
 * torque_tol > 0 adds a stop when max|m x H| / max|H| <= torque_tol (checked after each
 * iteration; the reference has no such stop, torque_tol <= 0 reproduces it exactly). */
typedef struct mafx_lbfgs_stats {
    int64_t iterations, energy_evals;
    double final_energy, final_torque_rel;
    int32_t stop_reason; /* 0 maxiter, 1 gradient, 2 triple test, 3 torque, 4 no descent,
                            5 stalled: five iterations in a row without an energy decrease
                            (checked only when torque_tol > 0) */
    int64_t finish_iterations; /* iterations run on the finish context (see below); included in
                                  iterations */
} mafx_lbfgs_stats;
/* verbose > 0 prints "f= E" at the start and, when the loop ends through a stop test other than
 * the gradient one or the iteration bound, the reference's closing "aa> iterations dE max|mxmxH|"
 * line (lbfgs_minimizer.cpp:181-192); verbose > 1 one "bb>" line per iteration (:142-146). */
int mafx_lbfgs_minimize(mafx_ctx* ctx, double* m_dev, double tolerance, int64_t max_iter,
                        int verbose, double torque_tol, double* E_out, mafx_lbfgs_stats* stats);
/* LBFGS_Minimizer::of_convergence_ (lbfgs_minimizer.hpp:18, lbfgs_minimizer.cpp:148-153): while a
 * file is set, every iteration of mafx_lbfgs_minimize on ctx writes the three stop ratios
 * (f_old - f) / (tol (1 + |f|)), max|s| / (sqrt(tol) (1 + max|m|)), max|g| / (tol^(1/3) (1 + |f|)),
 * tab-separated, one line per iteration.  The file is truncated when set; NULL closes it. */
int mafx_lbfgs_set_convergence_file(mafx_ctx* ctx, const char* path);
/* No counterpart in the reference (its minimizer runs in FP64 only).  With a finish context set
 * (a context of the same mesh and terms, normally the FP64 one of an FP32-field ctx), a
 * mafx_lbfgs_minimize on ctx with torque_tol > 0 that stalls (stop reason 5: the FP32 field's
 * energy rounding hides the decrease of an iteration) goes on from the same iterate on the finish
 * context, keeping the curvature pairs, H0 scaling and iteration count, until a stop test or
 * max_iter ends it.  NULL clears it.  The finish context must outlive its use here. */
int mafx_lbfgs_set_finish_context(mafx_ctx* ctx, mafx_ctx* finish);

/* ------------------------------------------------------------------ learned warm start */
/* No counterpart in the reference (magnum.af has no learned component).  A 4-level
 * convolutional encoder-decoder (3x3 convolutions 8->32->64->128->128 and back with skip
 * concatenation, BF16 weights, FP32 activations and accumulation) predicts a relaxed
 * magnetisation from m0 plus five parameter planes; LBFGS (mafx_lbfgs_minimize) then refines
 * it.  Layouts: inputs x [B][8][H][W] and outputs / targets [B][3][H][W], f32 device arrays, x
 * fastest; H and W multiples of 8.  Parameters: per layer W [Cout][Cin][3][3] then b [Cout],
 * concatenated in layer order (FP32 masters; the kernels use their BF16 rounding). */
typedef struct mafx_net mafx_net;
/* stream: a hipStream_t or NULL for the null stream.  Weights start He-normal (fixed seed). */
int mafx_net_create(void* stream, mafx_net** out);
int mafx_net_destroy(mafx_net* net);
int mafx_net_param_count(const mafx_net* net, int64_t* n);
int mafx_net_set_params(mafx_net* net, const float* host);
int mafx_net_get_params(const mafx_net* net, float* host);
/* m_dev [B][3][H][W] <- unit-vector prediction */
int mafx_net_forward(mafx_net* net, const float* x_dev, int B, int H, int W, float* m_dev);
/* loss = sum |m - target|^2 / (B H W); gradients (same layout as the parameters) to grad_host */
int mafx_net_gradients(mafx_net* net, const float* x_dev, const float* target_dev, int B, int H, int W,
                       float* grad_host, double* loss);
/* forward + backward + one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with learning rate lr */
int mafx_net_train_step(mafx_net* net, const float* x_dev, const float* target_dev, int B, int H, int W,
                        float lr, double* loss);
/* Warm start for a film (nz = 1, nx and ny multiples of 8): m_out_dev (SoA f64 [3][ny][nx],
 * renormalised in FP64) from m0_dev (same layout) and the five normalised parameter values
 * broadcast as input planes. */
int mafx_net_warm_start(mafx_net* net, const mafx_mesh* mesh, const double* m0_dev, const float planes[5],
                        double* m_out_dev);

/* ------------------------------------------------------------------ profiling */
/* When enabled, every kernel launch is bracketed by HIP events on the context stream and
 * its duration accumulated per kernel name. */
int mafx_prof_enable(mafx_ctx* ctx, int on);
int mafx_prof_reset(mafx_ctx* ctx);
/* Writes "name count total_ms bytes_per_launch\n" lines into buf (NUL-terminated). */
int mafx_prof_report(mafx_ctx* ctx, char* buf, int64_t buflen);

#ifdef __cplusplus
}
#endif
#endif /* MAFX_H */
