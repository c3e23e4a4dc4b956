// Measures the FP32 vector-ALU issue rate of the device: sixteen independent
// v_fma_f32 chains per lane, 16 waves requested per SIMD (the hardware keeps
// as many resident as fit), timed with HIP events over one launch after a
// warm-up launch; then the same with every fourth chain a v_rcp_f32. Prints
// one JSON line: wave-wide instructions issued per SIMD per second (the unit
// the SQ_INSTS_VALU counter counts in), the kernel times and the instruction
// count. The neighbour passes of the fluid step are bound by vector-ALU
// issue, not by HBM, and this is their ceiling in the same unit.
//
// Build: part of `make -C rust-wl_amd all` (no other source needed).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 16;     // independent FMA chains per lane
constexpr int kThreads = 256;   // four waves per workgroup
constexpr int kIters = 2000;    // loop trips per lane

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "valu_rate: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

// RCP = true: every fourth chain takes v_rcp_f32 instead of the FMA (the
// quarter-rate transcendental the correctly rounded division starts from).
template <bool RCP>
__global__ __launch_bounds__(kThreads) void k_fma_chains(const float *__restrict__ in, float *__restrict__ out, int n) {
    float a[kChains];
    const float b = in[threadIdx.x & 63u], c = in[64u + (threadIdx.x & 63u)];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = in[(threadIdx.x + k) & 127u];
    __builtin_amdgcn_s_waitcnt(0);  // the loads above are complete before the loop
    for (int r = 0; r < n; ++r) {
#pragma unroll
        for (int k = 0; k < kChains; ++k)
            a[k] = RCP && (k & 3) == 0 ? __builtin_amdgcn_rcpf(a[k]) : __builtin_fmaf(a[k], b, c);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * kThreads + threadIdx.x] = s;  // keeps every chain live
}

}  // namespace

int main() {
    int dev = 0;
    check(hipSetDevice(dev), "hipSetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    const int simds = prop.multiProcessorCount * 4;  // four SIMDs per compute unit
    const int blocks = simds * 4;                     // 16 waves per SIMD requested
    float host_in[128];
    for (int i = 0; i < 128; ++i) host_in[i] = 0.999f + 1e-5f * (float)i;
    float *in = nullptr, *out = nullptr;
    check(hipMalloc(&in, sizeof(host_in)), "hipMalloc");
    check(hipMalloc(&out, sizeof(float) * (size_t)kThreads * blocks), "hipMalloc");
    check(hipMemcpy(in, host_in, sizeof(host_in), hipMemcpyHostToDevice), "hipMemcpy");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    // One warm-up launch, then one timed launch; returns its milliseconds.
    auto timed = [&](void (*kern)(const float *, float *, int)) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), 0, 0, in, out, kIters);
        check(hipGetLastError(), "launch");
        check(hipEventRecord(e0, 0), "hipEventRecord");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), 0, 0, in, out, kIters);
        check(hipGetLastError(), "launch");
        check(hipEventRecord(e1, 0), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms;
    };
    const float ms = timed(k_fma_chains<false>);
    const float ms_rcp = timed(k_fma_chains<true>);
    // Wave-wide instructions of the loop (the few outside it are left out).
    const double waves = (double)blocks * (kThreads / 64);
    const double instr = waves * (double)kIters * kChains;
    const double rate = instr / (ms * 1e-3) / simds, rate_rcp = instr / (ms_rcp * 1e-3) / simds;
    std::printf("{\"fma_wave_instr_per_simd_s\": %.6e, \"kernel_ms\": %.4f, \"wave_instr\": %.6e, "
                "\"rcp_mix_wave_instr_per_simd_s\": %.6e, \"rcp_mix_kernel_ms\": %.4f, \"simds\": %d, "
                "\"device\": \"%s\"}\n",
                rate, ms, instr, rate_rcp, ms_rcp, simds, prop.gcnArchName);
    check(hipFree(in), "hipFree");
    check(hipFree(out), "hipFree");
    return 0;
}
