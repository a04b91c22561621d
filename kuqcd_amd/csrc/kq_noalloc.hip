// kq_mdaggm_noalloc: dst = sigma src - D D src (the CGM operator A, as kq_mdaggm)
// for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc). Everything kq_mdaggm takes from
// the allocator lives here in one static device array per GPU:
//
//   scratch = [ SoA source | SoA result | one partial sum per workgroup | sigma ]
//
// and the intermediate D src of the other parity is kept in the caller's dst
// buffer (exactly one SoA vector long) before the result is written over it in
// the host layout. The array's device address is looked up once per device with
// hipGetSymbolAddress (at the library's device set-up, or on the first call on
// another device); the name `kq_scratch` itself is never passed to a kernel.
// The work is queued on the library stream and the call returns without waiting
// (kq_synchronize waits for it).
// Calls on one GPU share the array: two calls that overlap on one GPU (from two
// host threads or two streams) race. Nothing in it has to start at zero: every
// byte a call reads was written earlier in the same call.
#include <string>

#include "kq_context.h"
#include "kq_layout.h"
#include "kuqcd_hip.h"

namespace kq {

// 64 MiB: the FP64 operator at 32^4 (two FP64 vectors of 25.2 MB and 8 KiB of partial sums)
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static_assert(KQ_NOALLOC_SCRATCH_BYTES == kNoAllocScratchBytes, "header and library disagree on the scratch size");

__device__ __align__(256) unsigned char kq_scratch[kNoAllocScratchBytes];

namespace {

__global__ void k_set_scalar(double *p, double v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};

size_t round_up(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

void *noalloc_scratch()
{
    int dev = 0;
    KQ_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("no-allocation scratch: device index out of range");
    if (g_scratch[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(kq_scratch));
        if (e != hipSuccess || p == nullptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(kq_scratch) failed: ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

static size_t noalloc_mdaggm_bytes(uint32_t Vh, size_t real_bytes)
{
    const size_t vec = round_up(2 * real_bytes * 3 * (size_t)Vh);
    return 2 * vec + round_up(sizeof(double) * dslash_nblk(Vh, kReproNT)) + 256;
}

template <class LF, class LL, typename R>
static void mdaggm_noalloc(LF fat, LL lng, const Geom &g, const R *src, R *dst, int parity, double sigma, hipStream_t st)
{
    using V2 = c2_t<R>;
    const size_t vec = round_up(sizeof(V2) * 3 * (size_t)g.Vh);
    unsigned char *base = static_cast<unsigned char *>(noalloc_scratch());
    V2 *s = reinterpret_cast<V2 *>(base);
    V2 *d = reinterpret_cast<V2 *>(base + vec);
    double *part = reinterpret_cast<double *>(base + 2 * vec);
    double *sig = reinterpret_cast<double *>(base + 2 * vec + round_up(sizeof(double) * dslash_nblk(g.Vh, kReproNT)));
    const size_t n3 = (size_t)3 * g.Vh;
    const Vec<R> sv{s, g.Vh, n3}, dv{d, g.Vh, n3};
    // the intermediate (other parity) in the caller's output buffer, SoA
    const Vec<R> tv{reinterpret_cast<V2 *>(dst), g.Vh, n3};
    vec_host_to_soa<R, R>(src, sv, 0, st);
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, sig, sigma);
    KQ_HIP_CHECK(hipGetLastError());
    DslashArgs<LF, LL, R> a{};
    a.g = g;
    a.fat = fat;
    a.lng = lng;
    a.nrhs = 1;
    a.dst_parity = 1 - parity;
    a.src = sv;
    a.dst = tv;
    launch_dslash(a, EPI_STORE, 1, kReproNT, st, 0);
    a.dst_parity = parity;
    a.src = tv;
    a.dst = dv;
    a.aux = sv;
    a.sigma = sig;
    a.partial = part;
    launch_dslash(a, EPI_AXMY_DOT, 1, kReproNT, st, 0);
    vec_soa_to_host<R, R>(dv, 0, dst, st);
}

// device memory of this process (not pageable or pinned host memory)
static bool is_device_memory(const void *p)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

}  // namespace kq

using namespace kq;

extern "C" KuQCDStatus kq_mdaggm_noalloc(const void *src, void *dst, int parity, double sigma)
{
    try {
        if (src == nullptr || dst == nullptr) {
            set_error("src or dst is nullptr");
            return KUQCD_ERR_NULL_POINTER;
        }
        if (parity != 0 && parity != 1) {
            set_error("parity must be 0 (even) or 1 (odd)");
            return KUQCD_ERR_INVALID_PARAM;
        }
        const int slot = active_slot();
        if (slot < 0) {
            set_error("no init gauge_smeared and gauge_naik");
            return KUQCD_ERR_UNKNOWN_TYPE;
        }
        Context &c = ctx();
        const Geom g = slot_geom(slot);
        const bool dbl = c.fat[slot].prec == KUQCD_DOUBLE;
        const size_t real = dbl ? sizeof(double) : sizeof(float);
        const size_t need = noalloc_mdaggm_bytes(g.Vh, real);
        if (need > kNoAllocScratchBytes) {
            set_error("kq_mdaggm_noalloc: the loaded lattice needs " + std::to_string(need) +
                      " bytes of scratch, above the fixed " + std::to_string(kNoAllocScratchBytes) +
                      " (64 MiB: up to 32^4 sites in double precision); use kq_mdaggm");
            return KUQCD_ERR_INVALID_PARAM;
        }
        const size_t bytes = real * 6 * (size_t)g.Vh;
        if (!is_device_memory(src) || !is_device_memory(dst)) {
            set_error("kq_mdaggm_noalloc: src and dst must be device memory (kq_device_alloc / hipMalloc)");
            return KUQCD_ERR_INVALID_PARAM;
        }
        const char *a = static_cast<const char *>(src), *b = static_cast<const char *>(dst);
        if (a < b + bytes && b < a + bytes) {
            set_error("kq_mdaggm_noalloc: src and dst overlap");
            return KUQCD_ERR_INVALID_PARAM;
        }
        // the same link formats as kq_mdaggm (with_fp64_links), so the two agree bitwise
        if (dbl)
            with_fp64_links(slot, [&](auto fat, auto lng) {
                mdaggm_noalloc<decltype(fat), decltype(lng), double>(fat, lng, g, static_cast<const double *>(src),
                                                                     static_cast<double *>(dst), parity, sigma,
                                                                     c.stream);
            });
        else
            with_fp32_links(slot, [&](auto fat, auto lng) {
                mdaggm_noalloc<decltype(fat), decltype(lng), float>(fat, lng, g, static_cast<const float *>(src),
                                                                    static_cast<float *>(dst), parity, sigma, c.stream);
            });
        return KUQCD_SUCCESS;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}
