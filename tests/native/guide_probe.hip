// guide_probe.hip - TEST INFRASTRUCTURE (not part of liblambent_hip.so).
//
// Runs the product's device functions of the guided mode
// (lambent_amd/csrc/dev_guide.h: the world hash grid, the Markov-chain update,
// kappa, the vMF pdf and sampling, the 21-candidate resampling step, the
// mixture pdf term, the screen-space candidate draw, the writer election) on
// the GPU over arrays of inputs, so tests can compare each one with the CPU
// oracle.  Every entry point copies host arrays in and results out and
// returns a hipError_t (0 = success).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "dev_guide.h"

using namespace lmb;

namespace {

constexpr int kBlock = 256;

struct Dev {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <class T>
    T* in(const T* host, uint64_t n) {
        T* d = nullptr;
        if (err != hipSuccess) return nullptr;
        err = hipMalloc(&d, (n ? n : 1) * sizeof(T));
        if (err != hipSuccess) return nullptr;
        ptrs.push_back(d);
        if (host && n) err = hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice);
        return d;
    }
    template <class T>
    T* out(uint64_t n) {
        return in<T>(nullptr, n);
    }
    template <class T>
    void back(T* host, const T* d, uint64_t n) {
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (err == hipSuccess && n) err = hipMemcpy(host, d, n * sizeof(T), hipMemcpyDeviceToHost);
    }
    ~Dev() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

uint32_t blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

__device__ V3 ld3(const float* p, uint64_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ Chain ldc(const uint32_t* w, uint64_t i) {
    const uint32_t* c = w + 8 * i;
    Chain s;
    s.ls = v3(__uint_as_float(c[0]), __uint_as_float(c[1]), __uint_as_float(c[2]));
    s.mean_cosine = __uint_as_float(c[3]);
    s.weight_sum = __uint_as_float(c[4]);
    s.num_samples = c[5];
    s.score = __uint_as_float(c[6]);
    return s;
}
__device__ void stc(uint32_t* w, uint64_t i, const Chain& s) {
    uint32_t* c = w + 8 * i;
    c[0] = __float_as_uint(s.ls.x);
    c[1] = __float_as_uint(s.ls.y);
    c[2] = __float_as_uint(s.ls.z);
    c[3] = __float_as_uint(s.mean_cosine);
    c[4] = __float_as_uint(s.weight_sum);
    c[5] = s.num_samples;
    c[6] = __float_as_uint(s.score);
    c[7] = 0u;
}

__global__ void k_grid_key(const float* pos, const float* nrm, const float* inc, uint64_t n,
                           V3 cam, uint32_t table, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const GridKey k = grid_key(ld3(pos, i), ld3(nrm, i), ld3(inc, i), cam, table);
    out[2 * i] = k.idx;
    out[2 * i + 1] = k.secondary;
}

__global__ void k_kappa(const uint32_t* chains, uint64_t n, float* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = get_kappa(ldc(chains, i));
}

__global__ void k_expm1(const float* x, uint64_t n, float* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    out[2 * i] = exp_minus_one(x[i]);
    out[2 * i + 1] = x_div_exp_minus_one(x[i]);
}

__global__ void k_vmf_pdf(const float* mean, const float* kappa, const float* dirs, uint64_t n,
                          float* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = vmf_pdf(ld3(mean, i), kappa[i], ld3(dirs, i));
}

__global__ void k_vmf_sample(const float* mean, const float* kappa, const uint32_t* seeds,
                             uint64_t n, float* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const V3 x = vmf_sample(ld3(mean, i), kappa[i], seeds[i]);
    out[3 * i] = x.x;
    out[3 * i + 1] = x.y;
    out[3 * i + 2] = x.z;
}

__global__ void k_update(const float* x1, const float* x2, const float* p_hat,
                         const uint32_t* seed, const uint32_t* not_vmf, const uint32_t* chains,
                         const float* conf, const float* sum_mc, uint64_t n, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    stc(out, i, update_markov_state(ld3(x1, i), ld3(x2, i), p_hat[i], seed[i], not_vmf[i] != 0u,
                                    ldc(chains, i), conf[i], sum_mc[i]));
}

__global__ void k_resample(const uint32_t* chains, const uint32_t* seeds, uint64_t n,
                           int32_t* sel, uint32_t* seed_after, float* sum_out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t seed = seeds[i];
    float sum_mc = 0.0f;
    int32_t s = -1;
    for (uint32_t k = 0; k < kNumResamples; ++k)
        if (resample_step(ldc(chains, i * kNumResamples + k), seed, sum_mc)) s = (int32_t)k;
    sel[i] = s;
    seed_after[i] = seed;
    sum_out[i] = sum_mc;
}

__global__ void k_mixture(const uint32_t* chains, const float* sum_mc, const float* hit,
                          const float* d, uint64_t n, float* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float pdf = 0.0f;
    if (sum_mc[i] > 0.0f)
        for (uint32_t k = 0; k < kNumResamples; ++k)
            pdf += mixture_term(ldc(chains, i * kNumResamples + k), ld3(hit, i), ld3(d, i), sum_mc[i]);
    out[i] = pdf;
}

constexpr uint32_t kCandWords = 3 + kScreenLastFrame;

__global__ void k_screen(const uint32_t* seeds, const uint32_t* xy, uint64_t n, uint32_t width,
                         uint32_t subgroup, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ScreenCands c = screen_candidates(seeds[i], xy[2 * i], xy[2 * i + 1], width, subgroup);
    uint32_t* o = out + kCandWords * i;
    o[0] = c.lane_pix;
    o[1] = c.lane_ok ? 1u : 0u;
    for (uint32_t k = 0; k < kScreenLastFrame; ++k) o[2 + k] = c.idx[k];
    o[2 + kScreenLastFrame] = c.seed_after;
}

// One wave: lane l holds key (keys[2l], keys[2l+1]) and takes part when bit l
// of `live` is set; out[l] = 1 when it is elected.
__global__ void k_first_writer(const uint32_t* keys, uint64_t live, uint32_t* out) {
    const uint32_t l = threadIdx.x;
    out[l] = 0u;
    if (!((live >> l) & 1ull)) return;
    GridKey k;
    k.idx = keys[2 * l];
    k.secondary = keys[2 * l + 1];
    out[l] = first_writer_of_key(k) ? 1u : 0u;
}

// Sequential world-grid operations by one thread (op codes as the oracle's
// grid hook: 0 load, 1 store, 2 radiance read, 3 radiance add); per op 12
// floats (pos, normal, incoming, radiance in), a force_old flag and a chain.
__global__ void k_grid_ops(WorldMarkovChain* cur, const WorldMarkovChain* old, uint32_t table,
                           V3 cam, const uint32_t* ops, const float* args, const uint32_t* force,
                           const uint32_t* chains_in, uint32_t nops, uint32_t* chains_out,
                           float* rad_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t i = 0; i < nops; ++i) {
        const float* a = args + 12 * i;
        const GridKey k = grid_key(v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), v3(a[6], a[7], a[8]),
                                   cam, table);
        Chain c = chain_zero();
        V3 r = v3s(0.0f);
        if (ops[i] == 0u) {
            c = grid_load(cur, old, k, force[i] != 0u);
        } else if (ops[i] == 1u) {
            grid_store(cur, k, ldc(chains_in, i));
        } else if (ops[i] == 2u) {
            r = grid_radiance(cur, old, k);
        } else if (ops[i] == 3u) {
            grid_add_radiance(cur, k, v3(a[9], a[10], a[11]));
        } else {
            // the fused add / load / update / store of a vertex's training,
            // with a fixed update of the loaded chain (the test restates it)
            grid_add_and_train(cur, old, k, v3(a[9], a[10], a[11]), force[i] != 0u,
                               [&](const Chain& wc) {
                                   c = wc;
                                   const Chain ci = ldc(chains_in, i);
                                   Chain st = wc;
                                   st.ls = ci.ls;
                                   st.mean_cosine = wc.mean_cosine * 0.5f;
                                   st.weight_sum = wc.weight_sum + 1.0f;
                                   st.num_samples = wc.num_samples + 3u;
                                   st.score = ci.score;
                                   return st;
                               });
        }
        stc(chains_out, i, c);
        rad_out[3 * i] = r.x;
        rad_out[3 * i + 1] = r.y;
        rad_out[3 * i + 2] = r.z;
    }
}

}  // namespace

extern "C" {

int probe_grid_key(const float* pos, const float* nrm, const float* inc, uint64_t n,
                   const float* cam, uint32_t table, uint32_t* out) {
    Dev d;
    float *dp = d.in(pos, 3 * n), *dn = d.in(nrm, 3 * n), *di = d.in(inc, 3 * n);
    uint32_t* dout = d.out<uint32_t>(2 * n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_grid_key, dim3(blocks(n)), dim3(kBlock), 0, 0, dp, dn, di, n,
                       V3{cam[0], cam[1], cam[2]}, table, dout);
    d.back(out, dout, 2 * n);
    return d.err;
}

int probe_kappa(const uint32_t* chains, uint64_t n, float* out) {
    Dev d;
    uint32_t* dc = d.in(chains, 8 * n);
    float* dout = d.out<float>(n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_kappa, dim3(blocks(n)), dim3(kBlock), 0, 0, dc, n, dout);
    d.back(out, dout, n);
    return d.err;
}

int probe_expm1(const float* x, uint64_t n, float* out) {
    Dev d;
    float* dx = d.in(x, n);
    float* dout = d.out<float>(2 * n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_expm1, dim3(blocks(n)), dim3(kBlock), 0, 0, dx, n, dout);
    d.back(out, dout, 2 * n);
    return d.err;
}

int probe_vmf_pdf(const float* mean, const float* kappa, const float* dirs, uint64_t n, float* out) {
    Dev d;
    float *dm = d.in(mean, 3 * n), *dk = d.in(kappa, n), *dd = d.in(dirs, 3 * n);
    float* dout = d.out<float>(n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_vmf_pdf, dim3(blocks(n)), dim3(kBlock), 0, 0, dm, dk, dd, n, dout);
    d.back(out, dout, n);
    return d.err;
}

int probe_vmf_sample(const float* mean, const float* kappa, const uint32_t* seeds, uint64_t n,
                     float* out) {
    Dev d;
    float *dm = d.in(mean, 3 * n), *dk = d.in(kappa, n);
    uint32_t* ds = d.in(seeds, n);
    float* dout = d.out<float>(3 * n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_vmf_sample, dim3(blocks(n)), dim3(kBlock), 0, 0, dm, dk, ds, n, dout);
    d.back(out, dout, 3 * n);
    return d.err;
}

int probe_update_state(const float* x1, const float* x2, const float* p_hat, const uint32_t* seed,
                       const uint32_t* not_vmf, const uint32_t* chains, const float* conf,
                       const float* sum_mc, uint64_t n, uint32_t* out) {
    Dev d;
    float *d1 = d.in(x1, 3 * n), *d2 = d.in(x2, 3 * n), *dp = d.in(p_hat, n);
    uint32_t *ds = d.in(seed, n), *dv = d.in(not_vmf, n), *dc = d.in(chains, 8 * n);
    float *dcf = d.in(conf, n), *dsm = d.in(sum_mc, n);
    uint32_t* dout = d.out<uint32_t>(8 * n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_update, dim3(blocks(n)), dim3(kBlock), 0, 0, d1, d2, dp, ds, dv, dc, dcf,
                       dsm, n, dout);
    d.back(out, dout, 8 * n);
    return d.err;
}

int probe_resample(const uint32_t* chains, const uint32_t* seeds, uint64_t n, int32_t* sel,
                   uint32_t* seed_after, float* sum_mc) {
    Dev d;
    uint32_t *dc = d.in(chains, 8 * kNumResamples * n), *ds = d.in(seeds, n);
    int32_t* dsel = d.out<int32_t>(n);
    uint32_t* dsa = d.out<uint32_t>(n);
    float* dsm = d.out<float>(n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_resample, dim3(blocks(n)), dim3(kBlock), 0, 0, dc, ds, n, dsel, dsa, dsm);
    d.back(sel, dsel, n);
    d.back(seed_after, dsa, n);
    d.back(sum_mc, dsm, n);
    return d.err;
}

int probe_mixture_pdf(const uint32_t* chains, const float* sum_mc, const float* hit,
                      const float* dirs, uint64_t n, float* out) {
    Dev d;
    uint32_t* dc = d.in(chains, 8 * kNumResamples * n);
    float *dsm = d.in(sum_mc, n), *dh = d.in(hit, 3 * n), *dd = d.in(dirs, 3 * n);
    float* dout = d.out<float>(n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_mixture, dim3(blocks(n)), dim3(kBlock), 0, 0, dc, dsm, dh, dd, n, dout);
    d.back(out, dout, n);
    return d.err;
}

int probe_screen_candidates(const uint32_t* seeds, const uint32_t* xy, uint64_t n, uint32_t width,
                            uint32_t subgroup, uint32_t* out) {
    if (subgroup == 0) return (int)hipErrorInvalidValue;
    Dev d;
    uint32_t *ds = d.in(seeds, n), *dxy = d.in(xy, 2 * n);
    uint32_t* dout = d.out<uint32_t>(kCandWords * n);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_screen, dim3(blocks(n)), dim3(kBlock), 0, 0, ds, dxy, n, width, subgroup,
                       dout);
    d.back(out, dout, kCandWords * n);
    return d.err;
}

int probe_first_writer(const uint32_t* keys, uint64_t live, uint32_t* out) {
    Dev d;
    uint32_t* dk = d.in(keys, 128);
    uint32_t* dout = d.out<uint32_t>(64);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_first_writer, dim3(1), dim3(64), 0, 0, dk, live, dout);
    d.back(out, dout, 64);
    return d.err;
}

// world / world_old: table cells of 56 bytes each (world updated in place).
int probe_grid_ops(void* world, const void* world_old, uint32_t table, const float* cam,
                   const uint32_t* ops, const float* args, const uint32_t* force,
                   const uint32_t* chains_in, uint32_t nops, uint32_t* chains_out, float* rad_out) {
    Dev d;
    auto* dw = d.in((const WorldMarkovChain*)world, table);
    auto* dwo = d.in((const WorldMarkovChain*)world_old, table);
    uint32_t *dop = d.in(ops, nops), *dfo = d.in(force, nops), *dci = d.in(chains_in, 8 * nops);
    float* da = d.in(args, 12 * (uint64_t)nops);
    uint32_t* dco = d.out<uint32_t>(8 * (uint64_t)nops);
    float* dro = d.out<float>(3 * (uint64_t)nops);
    if (d.err) return d.err;
    hipLaunchKernelGGL(k_grid_ops, dim3(1), dim3(64), 0, 0, dw, dwo, table,
                       V3{cam[0], cam[1], cam[2]}, dop, da, dfo, dci, nops, dco, dro);
    d.back((WorldMarkovChain*)world, dw, table);
    d.back(chains_out, dco, 8 * (uint64_t)nops);
    d.back(rad_out, dro, 3 * (uint64_t)nops);
    return d.err;
}

}  // extern "C"
