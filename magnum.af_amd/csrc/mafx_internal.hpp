// Internal declarations of libmafx: context layout, FFT plans, error handling, launch profiling.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mafx.h"

namespace mafx {

// magnumaf/include/magnumaf/constants.hpp:9-11
constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0 = 4e-7 * kPi;
constexpr double kGamma = 1.760859644e11 * kMu0;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define MAFX_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::mafx::Error(MAFX_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

inline void require(bool ok, const std::string& msg) {
    if (!ok) throw Error(MAFX_ERR_ARG, msg);
}

// ------------------------------------------------------------------ device buffers
struct DevBuf {
    void* p{nullptr};
    size_t bytes{0};
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b) MAFX_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ------------------------------------------------------------------ launch profiling
struct Prof {
    bool on{false};
    struct Entry {
        int64_t count{0};
        double ms{0};
        double bytes{0}; // algorithmic bytes per launch
    };
    std::map<std::string, Entry> entries;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<double> pending_bytes;
    std::vector<hipEvent_t> pool;
    hipEvent_t get_event();
    void flush(); // resolves pending events (synchronises)
    ~Prof();
};

// ------------------------------------------------------------------ FFT plan (one axis length)
struct FftPlan {
    int L{0};                 // transform length
    int nstages{0};
    int radix[16]{};          // radices in application order
    DevBuf tw64, tw32;        // exp(-2 pi i t / L), t in [0, L): double2 / float2
    void build(int len);      // throws if a prime factor is unsupported
};

// ------------------------------------------------------------------ terms
struct TermDev {
    int kind{0};
    int flags{0};
    double value{0};
    double axis[3]{0, 0, 1};
    double field[3]{0, 0, 0};
    DevBuf zee;       // 3N f64 when the external field is an array or a full-field callback
    DevBuf rkky, aex; // N f64
    DevBuf ridx;      // N u32 (optional)
    DevBuf rkw;       // 6N RKKY stencil weights per neighbour (x-, x+, y-, y+, z-, z+), in the stencils' type
    DevBuf pfield;    // N f64: D_constants (DMI) or Ku1_field (UNIAXIAL)
    DevBuf axfield;   // 3N f64: normalised Ku1_axis_field (UNIAXIAL)
    bool has_zee_array{false}, has_ridx{false};
    mafx_field_fn callback{nullptr}; // time-dependent external field
    void* callback_user{nullptr};
    bool callback_homogeneous{false};
    bool callback_host{false};        // m and the field pass through the pinned host buffers below
    std::vector<double> cb_m, cb_h;   // host staging of MAFX_TERM_HOST_CALLBACK (3N each)
};
// Uploads desc's parameters into t (construction and mafx_ctx_update_term).
void set_term(mafx_ctx* ctx, TermDev& t, const mafx_term_desc& desc);
// Fills t.rkw from t's RKKY arrays and the context's mesh, in the stencils' type of the precision.
void rkky_weights(mafx_ctx* ctx, TermDev& t);
inline size_t term_bytes(const TermDev& t) {
    return t.zee.bytes + t.rkky.bytes + t.aex.bytes + t.ridx.bytes + t.rkw.bytes + t.pfield.bytes + t.axfield.bytes;
}

struct Workspace; // precision-specific demag buffers

// The demag / dipole FFT convolution runs in FP64 (complex reference tensor) for MAFX_F64 and
// MAFX_F32_STENCILS, in FP32 (symmetric real tensor) for MAFX_F32_FIELD and MAFX_F32_DEMAG.
inline bool demag_f64(int precision) { return precision == MAFX_F64 || precision == MAFX_F32_STENCILS; }
inline bool stencils_f32(int precision) { return precision == MAFX_F32_FIELD || precision == MAFX_F32_STENCILS; }
// FP32 demag on the complex half spectrum of the whole tensor (MAFX_TERM_FULL_TENSOR) instead of
// the real symmetric quarter
inline bool demag_full_c32(int precision, int demag_flags) {
    return !demag_f64(precision) && (demag_flags & MAFX_TERM_FULL_TENSOR) != 0;
}
inline bool valid_precision(int p) {
    return p == MAFX_F64 || p == MAFX_F32_FIELD || p == MAFX_F32_DEMAG || p == MAFX_F32_STENCILS;
}

} // namespace mafx

struct mafx_ctx {
    mafx_mesh mesh{};
    int64_t N{0};
    double Ms{0};
    double time{0};           // State::t of the next evaluation (time-dependent external fields)
    bool has_ms_field{false};
    mafx::DevBuf ms_field; // N f64
    int precision{MAFX_F64};
    int num_cus{256};         // compute units of the current device (persistent grids)
    // RK stage inputs are formed in the per-cell pass from this many cells on (below it the
    // pass has too few lanes to hide the extra loads); MAFX_RK_FUSE_MIN_CELLS overrides it
    int64_t rk_fuse_min_cells{int64_t{1} << 20};
    // MAFX_F32_FIELD integrators keep their stage derivatives k_j in FP32 (MAFX_RK_K64=1: FP64)
    bool rk_k32{true};
    std::vector<mafx::TermDev> terms;
    int demag_term{-1};       // index of the (micro or atomistic) FFT term, -1 if none
    bool atom_dipole{false};
    int demag_flags{0};       // MAFX_TERM_CACHING / MAFX_TERM_VERBOSE of the FFT term
    int tensor_source{0};     // 0 no tensor, 1 computed, 2 read from the on-disk cache
    // FFT geometry: expanded lengths (demag_field.cpp:14-16) and half-spectrum width
    int64_t n0{0}, n1{0}, n2{0}, hx{0};
    int64_t nzs{0};           // z planes of the spectrum workspace: nz when the z stage is fused, else n2
    int64_t hxs{0};           // padded row stride of spectra (>= hx, multiple of 16)
    int64_t n1h{0}, n2h{0};   // n1/2 + 1, n2/2 + 1: ky, kz extent of the symmetric FP32 tensor
    mafx::FftPlan px, py, pz; // lengths n0, n1, n2 (px, py also supply the packing / split twiddles)
    mafx::FftPlan pxh, pyh;   // lengths nx, ny: the half-length x and split y transforms
    // FP64 path: [6][n2] planes of n1 hxs complex f64, the transformed reference tensor as it is.
    // FP32 path: [6][n2h] planes of n1h hxs real f32, the mirror-symmetric spectrum (fft.hip).
    // Planes store columns in tiles of 4 (tcol, fft_device.hpp).
    mafx::DevBuf tensor;
    int64_t res_off{0};       // MAFX_TERM_FULL_TENSOR: byte offset of the FP16 residual in `tensor`
    float res_inv[6]{};       // ... and its per-component 1 / scale
    mafx::DevBuf spec;        // [3][nzs] planes of n1 hxs complex values, column tiles
    mafx::DevBuf hdem;        // [3][N] FFT-convolution field (f64 or f32) between the x inverse and the cell pass
    mafx::DevBuf hbuf;        // 3N f64 scratch (field output for helpers)
    mafx::DevBuf partials;    // reduction partials (f64)
    mafx::DevBuf scal;        // small device scalars
    // LBFGS two-loop step whose final reduction (its dot product, alpha_i and coefficient) is
    // still pending in partials: the next two-loop pass forms it, or lbfgs_flush launches it
    struct LbfgsPending {
        int mode{-1}; // -1 none; 0 dot only; 1 alpha_i = rho d, coef = -alpha_i; 2 coef = alpha_i - rho d
        const double* part{nullptr};
        int nb{0};
        double rho{0};
        double* alpha_i{nullptr};
        double* coef{nullptr};
        double* dot{nullptr};
    } lbfgs_pend;
    double* host_scal{nullptr}; // pinned host scalars (hipHostMalloc)
    int64_t scal_nblocks{0};  // partial-sum blocks of the last evaluation
    int scal_nterms{0};       // terms selected in the last evaluation
    double last_max_torque{0}, last_max_h{0}; // max|m x H|, max|H| of the last gradient evaluation
    mafx::DevBuf vtmp;        // 3N f64 scratch for vector helpers
    hipStream_t stream{nullptr};
    std::FILE* lbfgs_convergence{nullptr}; // LBFGS_Minimizer::of_convergence_ (mafx_lbfgs_set_convergence_file)
    mafx_ctx* lbfgs_finish{nullptr}; // context an LBFGS run moves to when it stalls (mafx_lbfgs_set_finish_context)
    mafx::Prof prof;
    int64_t device_bytes{0};
    ~mafx_ctx();
};

namespace mafx {

// ------------------------------------------------------------------ host tensor setup (newell_host.cpp)
// Real-space Newell tensor [6][n2][n1][n0], reference operation order (demag_field.cpp:20-142).
void newell_tensor_host(const mafx_mesh& mesh, double* N, int nthreads);
// Its non-negative octant [6][nz][ny][nx] in the same order, and the octant's extension by parity to
// the expanded grid [6][n2][n1][n0] (unpaired planes zero), the FP32 path's real-space tensor.
void newell_octant_host(const mafx_mesh& mesh, double* O, int nthreads);
void newell_octant_extend(const mafx_mesh& mesh, const double* O, double* N);
// Point-dipole tensor of AtomisticDipoleDipoleField (atomistic_dipole_dipole_field.cpp:55-113), x 1/(4 pi).
void dipole_tensor_host(const mafx_mesh& mesh, double* N);

// ------------------------------------------------------------------ field evaluation (field.hip)
// stage vectors k_j the per-cell pass can combine into the next stage input (schemes of <= 8 stages)
constexpr int kMaxRkFused = 7;
enum class OutMode : int {
    H = 0,      // write H (f64 or f32)
    LLG = 1,    // write k = dt * LLG(alpha, m, H) (f64)
    GRAD = 2,   // write g = m x (m x H) (f64) and per-term energy partials
    ENERGY = 3  // per-term energy partials only
};
struct EvalArgs {
    const double* m{nullptr};
    uint32_t mask{~0u};
    OutMode mode{OutMode::H};
    void* out{nullptr};
    bool out_f32{false};
    double alpha{0}, dt{0};
    const double* alpha_field{nullptr}; // N f64 cell-wise damping, or null for `alpha`
    bool dissipation_only{false};
    bool want_energy{false};
    // OutMode::LLG: with rk_out set, the same pass also writes the next RK stage's input
    // rk_out = rk_base + sum_{j < rk_n} rk_c[j] rk_k[j] + rk_c[rk_n] k, k the output, in the order
    // and rounding of the separate stage combination.  With k_f32 the output k and the rk_k are
    // float32 (the sum is FP64, with this pass's k rounded to float32 as it is stored).
    bool k_f32{false};
    const double* rk_base{nullptr};
    const void* rk_k[kMaxRkFused]{};
    double rk_c[kMaxRkFused + 1]{};
    int rk_n{0};
    double* rk_out{nullptr};
};
// Runs the selected terms on m; the energy partial sums (one per term) land in ctx->partials and
// are reduced by finish_energy.
void evaluate(mafx_ctx* ctx, const EvalArgs& a);
void finish_energy(mafx_ctx* ctx, uint32_t mask, double* E_terms, double* E_total);
// The two halves of finish_energy: the reduction into ctx->scal[0, slots) (returns the slot count),
// and the term energies from those slots once they are in ctx->host_scal.
int finish_energy_launch(mafx_ctx* ctx);
void finish_energy_combine(mafx_ctx* ctx, uint32_t mask, double* E_terms, double* E_total);
void build_tensor(mafx_ctx* ctx);

// ------------------------------------------------------------------ vector kernels (vec.hip)
double dot(mafx_ctx* ctx, const double* a, const double* b, int64_t n);
double max_abs(mafx_ctx* ctx, const double* a, int64_t n);
void normalize(mafx_ctx* ctx, double* m, bool zero_safe);
void mean_m(mafx_ctx* ctx, const double* m, double out[3]);
void llg_rhs(mafx_ctx* ctx, double alpha, double dt, const double* m, const double* h, double* k);
// out = base + sum_j c_j v_j (j < n, n <= 12), evaluated left to right in FP64; the v_j are
// float32 when v_f32 (each converted to double as loaded)
void lincomb(mafx_ctx* ctx, double* out, const double* base, int n, const void* const* v, const double* c,
             int64_t len, bool v_f32 = false);
// x = normalize(a + s * b) with the zero-safe normalisation (lbfgs_minimizer.cpp:277-278)
void line_point(mafx_ctx* ctx, double* x, const double* a, double s, const double* b);

// ------------------------------------------------------------------ profiling helper
struct ProfScope {
    mafx_ctx* ctx;
    const char* name;
    double bytes;
    hipEvent_t b{nullptr};
    ProfScope(mafx_ctx* c, const char* n, double by);
    ~ProfScope();
};

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

} // namespace mafx
