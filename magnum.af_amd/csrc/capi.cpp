// extern "C" entry points of libmafx (declared in include/mafx.h): context construction, term
// upload, field / energy evaluation, vector helpers, the LLG integrator, LBFGS and profiling.
// No exception crosses the boundary: each entry point returns a status and records the message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fft_kernels.hpp"
#include "mafx_internal.hpp"
#include "tensor_cache.hpp"
#include "mafx_llg.hpp"
#include "net.hpp"
#include "vec_kernels.hpp"

namespace mafx {

namespace {
thread_local std::string g_last_error;

template <class F> int guarded(F&& f) {
    try {
        f();
        return MAFX_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return MAFX_ERR_DEVICE;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MAFX_ERR_DEVICE;
    }
}

// dmi_field.cpp:43-47, util.hpp:62-69, atomistic_dmi_field.cpp:6-9: v / sqrt(v0^2 + v1^2 + v2^2)
void normalize_axis(const double in[3], double out[3]) {
    const double n = std::sqrt(std::pow(in[0], 2) + std::pow(in[1], 2) + std::pow(in[2], 2));
    require(n > 0, "term axis must be non-zero");
    for (int c = 0; c < 3; ++c) out[c] = in[c] / n;
}

void upload(DevBuf& b, const void* host, size_t bytes) {
    b.alloc(bytes);
    MAFX_HIP(hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice));
}

} // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

hipEvent_t Prof::get_event() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    MAFX_HIP(hipEventCreate(&e));
    return e;
}

void Prof::flush() {
    for (size_t i = 0; i < pending.size(); ++i) {
        auto& p = pending[i];
        MAFX_HIP(hipEventSynchronize(p.second.second));
        float ms = 0;
        MAFX_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        Entry& e = entries[p.first];
        e.count++;
        e.ms += ms;
        e.bytes = pending_bytes[i];
        pool.push_back(p.second.first);
        pool.push_back(p.second.second);
    }
    pending.clear();
    pending_bytes.clear();
}

Prof::~Prof() {
    for (auto& p : pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(mafx_ctx* c, const char* n, double by) : ctx(c), name(n), bytes(by) {
    if (ctx->prof.on) {
        b = ctx->prof.get_event();
        MAFX_HIP(hipEventRecord(b, ctx->stream));
    }
}

ProfScope::~ProfScope() {
    if (!b) return;
    hipEvent_t e = nullptr;
    try {
        e = ctx->prof.get_event();
    } catch (...) {
        return;
    }
    if (hipEventRecord(e, ctx->stream) != hipSuccess) return;
    ctx->prof.pending.push_back({name, {b, e}});
    ctx->prof.pending_bytes.push_back(bytes);
    if (ctx->prof.pending.size() > 4096) {
        try {
            ctx->prof.flush();
        } catch (...) {
        }
    }
}

// Parameters of one term: scalars, normalised axes, and the arrays uploaded to the device.
void set_term(mafx_ctx* ctx, TermDev& t, const mafx_term_desc& d) {
    const int64_t N = ctx->N;
    require(t.kind == 0 || t.kind == d.kind, "a term keeps its kind");
    t.kind = d.kind;
    t.flags = d.flags;
    t.value = d.value;
    for (int c = 0; c < 3; ++c) t.field[c] = d.field[c];
    t.callback = nullptr;
    t.callback_user = nullptr;
    t.callback_homogeneous = false;
    t.callback_host = false;
    switch (d.kind) {
    case MAFX_DEMAG:
    case MAFX_ATOM_DIPOLE:
    case MAFX_EXCHANGE:
    case MAFX_ATOM_EXCHANGE: break;
    case MAFX_DMI:
    case MAFX_UNIAXIAL:
    case MAFX_ATOM_DMI:
    case MAFX_ATOM_UNIAXIAL:
        normalize_axis(d.axis, t.axis);
        t.pfield.release();
        t.axfield.release();
        if ((d.kind == MAFX_DMI || d.kind == MAFX_UNIAXIAL) && d.array0) upload(t.pfield, d.array0, sizeof(double) * N);
        if (d.kind == MAFX_UNIAXIAL && d.array1) {
            // util::normalize (util.cpp): every cell's axis divided by sqrt(a0 a0 + a1 a1 + a2 a2)
            std::vector<double> ax(d.array1, d.array1 + 3 * N);
            for (int64_t i = 0; i < N; ++i) {
                const double a0 = ax[i], a1 = ax[N + i], a2 = ax[2 * N + i];
                const double n = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
                ax[i] = a0 / n;
                ax[N + i] = a1 / n;
                ax[2 * N + i] = a2 / n;
            }
            upload(t.axfield, ax.data(), sizeof(double) * 3 * N);
        }
        break;
    case MAFX_ZEEMAN:
    case MAFX_ATOM_ZEEMAN:
        t.has_zee_array = false;
        if (d.callback) {
            t.callback = d.callback;
            t.callback_user = d.callback_user;
            t.callback_homogeneous = (d.flags & MAFX_TERM_HOMOGENEOUS_CALLBACK) != 0;
            t.callback_host = !t.callback_homogeneous && (d.flags & MAFX_TERM_HOST_CALLBACK) != 0;
            if (t.callback_host) {
                t.cb_m.resize(3 * N);
                t.cb_h.assign(3 * N, 0.0);
            }
            if (!t.callback_homogeneous) {
                if (t.zee.bytes != sizeof(double) * 3 * N) t.zee.alloc(sizeof(double) * 3 * N);
                t.has_zee_array = true;
            } else {
                t.zee.release();
            }
        } else if (d.array0) {
            upload(t.zee, d.array0, sizeof(double) * 3 * N);
            t.has_zee_array = true;
        } else {
            t.zee.release();
        }
        break;
    case MAFX_RKKY:
        require(d.array0 && d.array1, "RKKY needs rkky values (array0) and exchange values (array1)");
        upload(t.rkky, d.array0, sizeof(double) * N);
        upload(t.aex, d.array1, sizeof(double) * N);
        t.has_ridx = false;
        t.ridx.release();
        if (d.index0) {
            upload(t.ridx, d.index0, sizeof(uint32_t) * N);
            t.has_ridx = true;
        }
        rkky_weights(ctx, t);
        break;
    default: throw Error(MAFX_ERR_ARG, "unknown term kind " + std::to_string(d.kind));
    }
}

} // namespace mafx

mafx_ctx::~mafx_ctx() {
    if (host_scal) (void)hipHostFree(host_scal);
    if (lbfgs_convergence) std::fclose(lbfgs_convergence);
}

using namespace mafx;

extern "C" {

const char* mafx_last_error(void) { return g_last_error.c_str(); }

int mafx_device_count(int* count) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int mafx_demag_cache_name(const mafx_mesh* mesh, int precision, int term_flags, int dipole, char* buf,
                          int64_t buf_len) {
    return guarded([&] {
        require(mesh && buf && buf_len > 0, "mafx_demag_cache_name: null argument");
        require(valid_precision(precision), "unknown precision");
        TensorCacheKey k{2 * mesh->nx, 2 * mesh->ny, mesh->nz == 1 ? 1 : 2 * mesh->nz, (mesh->nx + 1 + 15) / 16 * 16,
                         tensor_cache_layout(demag_f64(precision) ? MAFX_F64 : MAFX_F32_FIELD, term_flags),
                         dipole ? 1 : 0, mesh->dx, mesh->dy, mesh->dz};
        const std::string s = tensor_cache_name(k);
        require(static_cast<int64_t>(s.size()) < buf_len, "mafx_demag_cache_name: buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int mafx_ctx_tensor_source(const mafx_ctx* ctx, int* source) {
    return guarded([&] {
        require(ctx && source, "mafx_ctx_tensor_source: null argument");
        *source = ctx->tensor_source;
    });
}

int mafx_demag_calc_N(const mafx_mesh* mesh, double* N_host, int nthreads) {
    return guarded([&] {
        require(mesh && N_host, "null argument");
        require(mesh->nx > 0 && mesh->ny > 0 && mesh->nz > 0, "mesh sizes must be positive");
        newell_tensor_host(*mesh, N_host, nthreads);
    });
}

int mafx_ctx_create(const mafx_mesh* mesh, double Ms, const double* Ms_field_host, const mafx_term_desc* terms,
                    int nterms, int precision, mafx_ctx** out) {
    return guarded([&] {
        require(mesh && out, "null argument");
        require(mesh->nx > 0 && mesh->ny > 0 && mesh->nz > 0, "mesh sizes must be positive");
        require(mesh->dx > 0 && mesh->dy > 0 && mesh->dz > 0, "cell sizes must be positive");
        require(nterms >= 1 && nterms <= 8, "between 1 and 8 terms per context");
        require(valid_precision(precision), "unknown precision");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw Error(MAFX_ERR_DEVICE, "no HIP device visible: the MI355X path needs a GPU");
        auto ctx = std::make_unique<mafx_ctx>();
        int dev = 0, cus = 0;
        MAFX_HIP(hipGetDevice(&dev));
        MAFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus > 0) ctx->num_cus = cus;
        // MAFX_RK_FUSE_MIN_CELLS: mesh size from which RK stage inputs are formed in the per-cell
        // pass (tests compare both forms on small meshes); MAFX_RK_K64=1 keeps the FP32-field
        // integrator's stage derivatives in FP64
        if (const char* e = std::getenv("MAFX_RK_FUSE_MIN_CELLS")) ctx->rk_fuse_min_cells = std::atoll(e);
        if (const char* e = std::getenv("MAFX_RK_K64")) ctx->rk_k32 = std::atoi(e) == 0;
        // MAFX_NUM_CUS overrides the CU count that sizes the persistent grids (tuning, and tests
        // that make each resident group take many tiles on small meshes)
        if (const char* e = std::getenv("MAFX_NUM_CUS")) {
            const int v = std::atoi(e);
            if (v > 0) ctx->num_cus = v;
        }
        ctx->mesh = *mesh;
        ctx->N = mesh->nx * mesh->ny * mesh->nz;
        ctx->Ms = Ms;
        ctx->precision = precision;
        const int64_t N = ctx->N;
        if (Ms_field_host) {
            upload(ctx->ms_field, Ms_field_host, sizeof(double) * N);
            ctx->has_ms_field = true;
        }
        for (int i = 0; i < nterms; ++i) {
            const mafx_term_desc& d = terms[i];
            TermDev t;
            if (d.kind == MAFX_DEMAG || d.kind == MAFX_ATOM_DIPOLE) {
                require(ctx->demag_term < 0, "at most one demag / dipole term per context");
                ctx->demag_term = i;
                ctx->atom_dipole = d.kind == MAFX_ATOM_DIPOLE;
                ctx->demag_flags = d.flags;
            }
            set_term(ctx.get(), t, d);
            ctx->terms.push_back(std::move(t));
        }
        // partial-sum blocks: epilogue row groups, per-cell tiles (row tiles of 256 cells or
        // memory order), vector reductions
        int64_t nblk = std::max<int64_t>(mesh->ny * mesh->nz * ((mesh->nx + 255) / 256), (N + 255) / 256);
        nblk = std::max<int64_t>(nblk, reduction_partials_needed(3 * N));
        ctx->partials.alloc(sizeof(double) * 10 * nblk);
        ctx->scal.alloc(sizeof(double) * 64);
        MAFX_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_scal), sizeof(double) * 64, hipHostMallocDefault));
        ctx->hbuf.alloc(sizeof(double) * 3 * N);
        if (ctx->demag_term >= 0) {
            ctx->n0 = 2 * mesh->nx;
            ctx->n1 = 2 * mesh->ny;
            ctx->n2 = mesh->nz == 1 ? 1 : 2 * mesh->nz;
            ctx->hx = ctx->n0 / 2 + 1;
            ctx->hxs = (ctx->hx + 15) / 16 * 16; // whole column tiles; planes start on 512 B boundaries
            ctx->px.build(static_cast<int>(ctx->n0));  // twiddles of the length-2 nx packing
            ctx->pxh.build(static_cast<int>(mesh->nx)); // the half-length x transform
            ctx->py.build(static_cast<int>(ctx->n1));
            ctx->pyh.build(static_cast<int>(mesh->ny)); // the split y transform
            if (ctx->n2 > 1) ctx->pz.build(static_cast<int>(ctx->n2));
            // the z transform is fused into the tensor product for 2 <= nz <= 8 (nz planes suffice)
            ctx->nzs = (mesh->nz >= 2 && mesh->nz <= 8) ? mesh->nz : ctx->n2;
            const size_t csz = demag_f64(precision) ? sizeof(double2) : sizeof(float2);
            ctx->spec.alloc(csz * 3 * ctx->nzs * ctx->n1 * ctx->hxs);
            MAFX_HIP(hipMemset(ctx->spec.p, 0, ctx->spec.bytes));
            ctx->hdem.alloc((demag_f64(precision) ? sizeof(double) : sizeof(float)) * 3 * N);
            build_tensor(ctx.get());
        }
        ctx->device_bytes = static_cast<int64_t>(ctx->ms_field.bytes + ctx->partials.bytes + ctx->scal.bytes +
                                                 ctx->hbuf.bytes + ctx->spec.bytes + ctx->tensor.bytes +
                                                 ctx->hdem.bytes);
        for (auto& t : ctx->terms)
            ctx->device_bytes += static_cast<int64_t>(term_bytes(t));
        *out = ctx.release();
    });
}

int mafx_ctx_destroy(mafx_ctx* ctx) {
    return guarded([&] {
        if (ctx) {
            (void)hipStreamSynchronize(ctx->stream);
            delete ctx;
        }
    });
}

int mafx_ctx_set_stream(mafx_ctx* ctx, void* stream) {
    return guarded([&] {
        require(ctx, "null context");
        ctx->stream = static_cast<hipStream_t>(stream);
    });
}

int mafx_ctx_sync(mafx_ctx* ctx) {
    return guarded([&] { MAFX_HIP(hipStreamSynchronize(ctx->stream)); });
}

int mafx_ctx_device_bytes(const mafx_ctx* ctx, int64_t* bytes) {
    return guarded([&] { *bytes = ctx->device_bytes; });
}

int mafx_ctx_set_time(mafx_ctx* ctx, double t) {
    return guarded([&] {
        require(ctx, "null context");
        ctx->time = t;
    });
}

int mafx_ctx_update_term(mafx_ctx* ctx, int index, const mafx_term_desc* desc) {
    return guarded([&] {
        require(ctx && desc, "null argument");
        require(index >= 0 && index < static_cast<int>(ctx->terms.size()), "term index out of range");
        require(desc->kind == ctx->terms[index].kind, "mafx_ctx_update_term: the term kind cannot change");
        // kernels of earlier calls may still read the arrays being replaced
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        TermDev& t = ctx->terms[index];
        const int64_t before = static_cast<int64_t>(term_bytes(t));
        set_term(ctx, t, *desc);
        ctx->device_bytes += static_cast<int64_t>(term_bytes(t)) - before;
    });
}

int mafx_ctx_set_material(mafx_ctx* ctx, double Ms, const double* Ms_field_host) {
    return guarded([&] {
        require(ctx, "null context");
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        ctx->device_bytes -= static_cast<int64_t>(ctx->ms_field.bytes);
        ctx->Ms = Ms;
        if (Ms_field_host) {
            upload(ctx->ms_field, Ms_field_host, sizeof(double) * ctx->N);
            ctx->has_ms_field = true;
        } else {
            ctx->ms_field.release();
            ctx->has_ms_field = false;
        }
        ctx->device_bytes += static_cast<int64_t>(ctx->ms_field.bytes);
    });
}

int mafx_heff(mafx_ctx* ctx, const double* m_dev, uint32_t term_mask, void* h_dev, int h_is_f32) {
    return guarded([&] {
        require(ctx && m_dev && h_dev, "null argument");
        EvalArgs a;
        a.m = m_dev;
        a.mask = term_mask;
        a.mode = OutMode::H;
        a.out = h_dev;
        a.out_f32 = h_is_f32 != 0;
        evaluate(ctx, a);
    });
}

int mafx_energy(mafx_ctx* ctx, const double* m_dev, uint32_t term_mask, double* E_terms_host, double* E_total_host) {
    return guarded([&] {
        require(ctx && m_dev, "null argument");
        EvalArgs a;
        a.m = m_dev;
        a.mask = term_mask;
        a.mode = OutMode::ENERGY;
        a.want_energy = true;
        evaluate(ctx, a);
        finish_energy(ctx, term_mask, E_terms_host, E_total_host);
    });
}

int mafx_llg_rhs(mafx_ctx* ctx, double alpha, double dt, const double* m_dev, const double* h_dev, double* k_dev) {
    return guarded([&] { llg_rhs(ctx, alpha, dt, m_dev, h_dev, k_dev); });
}

int mafx_dot(mafx_ctx* ctx, const double* a_dev, const double* b_dev, double* out_host) {
    return guarded([&] { *out_host = dot(ctx, a_dev, b_dev, 3 * ctx->N); });
}

int mafx_max_abs(mafx_ctx* ctx, const double* a_dev, double* out_host) {
    return guarded([&] { *out_host = max_abs(ctx, a_dev, 3 * ctx->N); });
}

int mafx_normalize(mafx_ctx* ctx, double* m_dev, int zero_safe) {
    return guarded([&] { normalize(ctx, m_dev, zero_safe != 0); });
}

int mafx_mean_m(mafx_ctx* ctx, const double* m_dev, double out_host[3]) {
    return guarded([&] { mean_m(ctx, m_dev, out_host); });
}

int mafx_mean_m_noalloc(const double* m_dev, int64_t ncells, double* out_dev, void* stream) {
    return guarded([&] {
        require(m_dev != nullptr && out_dev != nullptr, "mafx_mean_m_noalloc: null pointer");
        require(ncells >= 1 && ncells <= (int64_t(1) << 60), "mafx_mean_m_noalloc: ncells must be in [1, 2^60]");
        mean_m_noalloc(m_dev, ncells, out_dev, static_cast<hipStream_t>(stream));
    });
}

int mafx_llg_create(mafx_ctx* ctx, double alpha, const char* scheme, double hmin, double hmax, double atol,
                    double rtol, int dissipation_term_only, mafx_llg** out) {
    return guarded([&] {
        require(ctx && scheme && out, "null argument");
        const std::string s(scheme);
        require(s == "DP45" || s == "RKF45" || s == "BS23" || s == "BS45" || s == "DP78",
                "scheme must be one of \"RKF45\", \"DP45\", \"BS45\", \"DP78\", \"BS23\"");
        auto L = std::make_unique<mafx_llg>();
        L->ctx = ctx;
        L->alpha = alpha;
        L->scheme = s;
        L->dissipation_only = dissipation_term_only != 0;
        L->ctrl.hmin = hmin;
        L->ctrl.hmax = hmax;
        L->ctrl.atol = atol;
        L->ctrl.rtol = rtol;
        const size_t bytes = sizeof(double) * 3 * ctx->N;
        // the scheme's stages; at least RKF45's six for the dense-output back step
        const int ns = std::max(scheme_stages(s.c_str()), 6);
        for (int j = 1; j <= ns; ++j) L->k[j].alloc(bytes);
        L->mtmp.alloc(bytes);
        L->mtmp2.alloc(bytes);
        L->sumbk.alloc(bytes);
        *out = L.release();
    });
}

int mafx_llg_destroy(mafx_llg* llg) {
    return guarded([&] { delete llg; });
}

int mafx_llg_set_alpha(mafx_llg* llg, double alpha) {
    return guarded([&] { llg->alpha = alpha; });
}

int mafx_llg_set_alpha_field(mafx_llg* llg, const double* alpha_host) {
    return guarded([&] {
        require(llg, "null integrator");
        MAFX_HIP(hipStreamSynchronize(llg->ctx->stream));
        if (alpha_host) {
            upload(llg->alpha_field, alpha_host, sizeof(double) * llg->ctx->N);
            llg->has_alpha_field = true;
        } else {
            llg->alpha_field.release();
            llg->has_alpha_field = false;
        }
    });
}

int mafx_llg_step(mafx_llg* llg, double* m_dev, double* t_inout) {
    return guarded([&] {
        require(llg && m_dev && t_inout, "null argument");
        llg_step(llg, m_dev, t_inout);
    });
}

int mafx_llg_relax(mafx_llg* llg, double* m_dev, double* t_inout, double precision, int64_t eval_E,
                   int64_t iwritecout, int verbose, int64_t state_steps, int64_t* nsteps_done) {
    return guarded([&] {
        require(llg && m_dev && t_inout, "null argument");
        require(eval_E > 0, "eval_E must be positive");
        mafx_ctx* ctx = llg->ctx;
        auto energy = [&] {
            ctx->time = *t_inout;
            EvalArgs a;
            a.m = m_dev;
            a.mode = OutMode::ENERGY;
            a.want_energy = true;
            evaluate(ctx, a);
            double E = 0;
            finish_energy(ctx, ~0u, nullptr, &E);
            return E;
        };
        const double t0 = *t_inout;
        int64_t n = 0;
        double E_prev = 1e20;
        double E = energy();
        while (std::fabs((E_prev - E) / E_prev) > precision) {
            E_prev = E;
            for (int64_t i = 0; i < eval_E; ++i) {
                llg_step(llg, m_dev, t_inout);
                ++n;
            }
            E = energy();
            // llg_integrator.cpp:52: the State's step count decides the print cadence
            const int64_t steps = state_steps + n;
            if (verbose && iwritecout > 0 && steps % iwritecout == 0) {
                double mm[3];
                mean_m(ctx, m_dev, mm);
                std::printf("relaxing: step=%lld, rel_diff= %e, <mx>=%f, <my>=%f, <mz>=%f\n",
                            static_cast<long long>(steps), std::fabs((E_prev - E) / E_prev), mm[0], mm[1], mm[2]);
            }
        }
        if (verbose) std::printf("relaxed %e [s]. Current state.t= %e\n", *t_inout - t0, *t_inout);
        if (nsteps_done) *nsteps_done = n;
    });
}

int mafx_llg_run_until(mafx_llg* llg, double* m_dev, double* t_inout, double t_end, int64_t nsteps_max,
                       int64_t* nsteps_done) {
    return guarded([&] {
        int64_t n = 0;
        while (*t_inout < t_end && n < nsteps_max) {
            llg_step(llg, m_dev, t_inout);
            ++n;
        }
        if (nsteps_done) *nsteps_done = n;
    });
}

int mafx_llg_integrate_dense(mafx_llg* llg, double* m_dev, double* t_inout, double time_in_s, double write_every_dt,
                             double* out_host, int64_t max_rows, int64_t* rows_written) {
    return guarded([&] {
        require(write_every_dt > 0, "write_every_dt must be positive");
        const int64_t r = llg_integrate_dense(llg, m_dev, t_inout, time_in_s, write_every_dt, out_host, max_rows);
        if (rows_written) *rows_written = std::min(r, max_rows);
    });
}

int mafx_llg_get_stats(const mafx_llg* llg, mafx_llg_stats* out) {
    return guarded([&] {
        out->h = llg->h;
        out->h_last = llg->h_last;
        out->err_last = llg->err_last;
        out->accepted = llg->ctrl.counter_accepted;
        out->rejected = llg->ctrl.counter_reject;
        out->heff_evals = llg->heff_evals;
        out->counter_hmin = llg->ctrl.counter_hmin;
        out->counter_hmax = llg->ctrl.counter_hmax;
        out->counter_maxscale = llg->ctrl.counter_maxscale;
        out->counter_minscale = llg->ctrl.counter_minscale;
    });
}

int mafx_llg_set_controller_state(mafx_llg* llg, double h, double errold, int reject) {
    return guarded([&] {
        llg->h = h;
        llg->ctrl.errold = errold;
        llg->ctrl.reject = reject != 0;
    });
}

int mafx_llg_get_controller_state(const mafx_llg* llg, double* h, double* errold, int* reject) {
    return guarded([&] {
        *h = llg->h;
        *errold = llg->ctrl.errold;
        *reject = llg->ctrl.reject ? 1 : 0;
    });
}

int mafx_lbfgs_minimize(mafx_ctx* ctx, double* m_dev, double tolerance, int64_t max_iter, int verbose,
                        double torque_tol, double* E_out, mafx_lbfgs_stats* stats) {
    int rc = MAFX_OK;
    const int g = guarded([&] { rc = lbfgs_minimize(ctx, m_dev, tolerance, max_iter, verbose, torque_tol, E_out, stats); });
    return g != MAFX_OK ? g : rc;
}

int mafx_lbfgs_set_convergence_file(mafx_ctx* ctx, const char* path) {
    return guarded([&] {
        require(ctx, "null context");
        if (ctx->lbfgs_convergence) std::fclose(ctx->lbfgs_convergence);
        ctx->lbfgs_convergence = nullptr;
        if (path) {
            ctx->lbfgs_convergence = std::fopen(path, "w");
            require(ctx->lbfgs_convergence != nullptr, std::string("cannot open ") + path);
        }
    });
}

int mafx_lbfgs_set_finish_context(mafx_ctx* ctx, mafx_ctx* finish) {
    return guarded([&] {
        require(ctx, "null context");
        if (finish) {
            require(finish->N == ctx->N && finish->mesh.nx == ctx->mesh.nx && finish->mesh.ny == ctx->mesh.ny &&
                        finish->mesh.nz == ctx->mesh.nz,
                    "mafx_lbfgs_set_finish_context: the finish context has another mesh");
            require(finish->terms.size() == ctx->terms.size(),
                    "mafx_lbfgs_set_finish_context: the finish context has another term count");
        }
        ctx->lbfgs_finish = finish == ctx ? nullptr : finish;
    });
}

int mafx_net_create(void* stream, mafx_net** out) {
    return guarded([&] {
        require(out != nullptr, "mafx_net_create: null output");
        int n = 0;
        MAFX_HIP(hipGetDeviceCount(&n));
        require(n > 0, "no GPU");
        *out = reinterpret_cast<mafx_net*>(new Net(static_cast<hipStream_t>(stream)));
    });
}

int mafx_net_destroy(mafx_net* net) {
    return guarded([&] { delete reinterpret_cast<Net*>(net); });
}

int mafx_net_param_count(const mafx_net* net, int64_t* n) {
    return guarded([&] { *n = reinterpret_cast<const Net*>(net)->nparams; });
}

int mafx_net_set_params(mafx_net* net, const float* host) {
    return guarded([&] { reinterpret_cast<Net*>(net)->set_params(host); });
}

int mafx_net_get_params(const mafx_net* net, float* host) {
    return guarded([&] { reinterpret_cast<const Net*>(net)->get_params(host); });
}

int mafx_net_forward(mafx_net* net_, const float* x_dev, int B, int H, int W, float* m_dev) {
    return guarded([&] {
        Net* net = reinterpret_cast<Net*>(net_);
        net->ensure_shape(B, H, W);
        net->forward(x_dev);
        MAFX_HIP(hipMemcpyAsync(m_dev, net->output(), sizeof(float) * 3 * static_cast<int64_t>(B) * H * W,
                                hipMemcpyDeviceToDevice, net->stream));
        MAFX_HIP(hipStreamSynchronize(net->stream));
    });
}

int mafx_net_gradients(mafx_net* net_, const float* x_dev, const float* target_dev, int B, int H, int W,
                       float* grad_host, double* loss) {
    return guarded([&] {
        Net* net = reinterpret_cast<Net*>(net_);
        net->ensure_shape(B, H, W);
        net->forward(x_dev);
        *loss = net->backward(target_dev);
        MAFX_HIP(hipMemcpyAsync(grad_host, net->grad.p, sizeof(float) * net->nparams, hipMemcpyDeviceToHost,
                                net->stream));
        MAFX_HIP(hipStreamSynchronize(net->stream));
    });
}

int mafx_net_train_step(mafx_net* net_, const float* x_dev, const float* target_dev, int B, int H, int W, float lr,
                        double* loss) {
    return guarded([&] {
        Net* net = reinterpret_cast<Net*>(net_);
        net->ensure_shape(B, H, W);
        net->forward(x_dev);
        *loss = net->backward(target_dev);
        net->adam(lr, 0.9f, 0.999f, 1e-8f);
    });
}

int mafx_net_warm_start(mafx_net* net_, const mafx_mesh* mesh, const double* m0_dev, const float planes[5],
                        double* m_out_dev) {
    return guarded([&] {
        Net* net = reinterpret_cast<Net*>(net_);
        require(mesh->nz == 1, "warm start: films only (nz = 1)");
        const int nx = static_cast<int>(mesh->nx), ny = static_cast<int>(mesh->ny);
        net->ensure_shape(1, ny, nx);
        DevBuf x;
        x.alloc(sizeof(float) * kNetInputs * static_cast<int64_t>(nx) * ny);
        net_pack_input(net, m0_dev, nx, ny, planes, x.as<float>());
        net->forward(x.as<float>());
        net_unpack_output(net, net->output(), nx, ny, m_out_dev);
        MAFX_HIP(hipStreamSynchronize(net->stream));
    });
}

int mafx_prof_enable(mafx_ctx* ctx, int on) {
    return guarded([&] { ctx->prof.on = on != 0; });
}

int mafx_prof_reset(mafx_ctx* ctx) {
    return guarded([&] {
        ctx->prof.flush();
        ctx->prof.entries.clear();
    });
}

int mafx_prof_report(mafx_ctx* ctx, char* buf, int64_t buflen) {
    return guarded([&] {
        ctx->prof.flush();
        std::string s;
        char line[256];
        for (auto& kv : ctx->prof.entries) {
            std::snprintf(line, sizeof(line), "%s %lld %.6f %.0f\n", kv.first.c_str(),
                          static_cast<long long>(kv.second.count), kv.second.ms, kv.second.bytes);
            s += line;
        }
        require(buflen > static_cast<int64_t>(s.size()), "report buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

} // extern "C"
