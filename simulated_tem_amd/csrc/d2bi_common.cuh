// d2bi_common.cuh -- error state, device buffers and the library state, device
// atom sets, and the generic exclusive scan.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ errors

static char g_err[1024] = "";

int seterr(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return seterr(D2BI_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                           \
    } while (0)

#define CHK(call)                        \
    do {                                 \
        int r_ = (call);                 \
        if (r_ != D2BI_OK) return r_;    \
    } while (0)

// ------------------------------------------------------------------ device buffers

static uint64_t g_alloc_gen = 0;         // bumped on every device reallocation

// A fixed block of device memory handed out front to back (256-byte
// aligned pieces); buffers bound to an arena never call the allocator.
struct Arena {
    char* base = nullptr;
    size_t size = 0, off = 0;
};

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;                  // set: pieces of the arena instead of hipMalloc
    int ensure(size_t bytes)
    {
        if (bytes <= cap && p) return D2BI_OK;
        if (arena) {
            const size_t off = (arena->off + 255) & ~(size_t)255;
            if (off > arena->size || bytes > arena->size - off)
                return seterr(D2BI_ERR_CAPACITY,
                              "the no-allocation scratch (%zu bytes) cannot hold this condition: "
                              "%zu bytes in use, %zu more requested",
                              arena->size, arena->off, bytes);
            p = arena->base + off;
            cap = bytes;
            arena->off = off + bytes;
            return D2BI_OK;
        }
        g_alloc_gen++;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        const size_t want = alloc_size(bytes);       // headroom for slightly larger conditions
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return seterr(D2BI_ERR_HIP, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        cap = want;
        return D2BI_OK;
    }
    void release()
    {
        if (p && !arena) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    // bytes ensure(bytes) allocates: the request with its headroom
    static size_t alloc_size(size_t bytes)
    {
        const size_t want = std::max<size_t>(bytes, 256);
        return want + want / 8;
    }
    // net device bytes ensure(bytes) would add (it frees the old block first)
    double growth(size_t bytes) const
    {
        if (bytes <= cap && p) return 0.0;
        return (double)alloc_size(bytes) - (double)cap;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Workspace {
    // binning
    Buf x;                      // atoms (temporary upload)
    Buf mask, blk_cnt, blk_off;
    Buf pt_cell, pt_slot, pt_frac, pt_gdotr;
    Buf pt_j;                   // atom number of each point (FAST imaging, FP64 aperture sums)
    Buf cell_cnt, cell_start, cell_ids;
    Buf slot_id, slot_frac, slot_cs64, slot_p32, slot_sn32;
    Buf app_cs;                 // per-slot phases of the aperture points of one shared pass
    Buf scan_tmp;
    Buf blk_max;                // per-block largest cell count (link-cell sort)
    Buf guard_part;             // FAST guard re-test counters, one per 128-byte line
    Buf counters;               // misc device counters (guard count, max per cell)
    // voxel axis tables
    Buf ax_cell, ax_zi, ax_start, ax_ql;
    // fields and image
    Buf f64_re, f64_im, f64_rho, f32_field;
    Buf cnt, r2m;
    Buf img, aa;                // image; 2 pi s_g per precession point (FP64 integration)
    Buf tune_img;               // image written by timing launches of the tuner
    Buf blur_in, blur_tmp, blur_out, blur_w, blur_wy, blur_k;
    // pinned host staging
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    template <class F> void each(F f)
    {
        Buf* all[] = {&x, &mask, &blk_cnt, &blk_off, &pt_cell, &pt_slot, &pt_frac, &pt_gdotr, &pt_j,
                      &cell_cnt, &cell_start, &cell_ids, &slot_id, &slot_frac, &slot_cs64,
                      &slot_p32, &slot_sn32, &app_cs, &scan_tmp, &blk_max, &guard_part, &counters, &ax_cell, &ax_zi, &ax_start, &ax_ql, &f64_re,
                      &f64_im, &f64_rho, &f32_field, &cnt, &r2m, &img, &aa, &tune_img, &blur_in, &blur_tmp,
                      &blur_out, &blur_w, &blur_wy, &blur_k};
        for (Buf* b : all) f(*b);
    }
    // every buffer a piece of arena a (no device allocation from then on)
    void bind(Arena* a)
    {
        each([&](Buf& b) { b.arena = a; });
    }
    // device bytes held
    size_t bytes()
    {
        size_t n = 0;
        each([&](Buf& b) { n += b.cap; });
        return n;
    }
    void release_all()
    {
        each([](Buf& b) { b.release(); });
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_cap = 0;
    }
};

struct State {
    int device = -1;
    bool init = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    bool pinned_busy = false;       // ev[9] marks the last copy out of the pinned staging buffer
    Workspace ws;
};

static State g;

static int need_init()
{
    if (!g.init) return seterr(D2BI_ERR_NOINIT, "d2bi_gpu_init has not been called");
    return D2BI_OK;
}

// The pinned staging buffer: a caller that leaves an asynchronous copy from
// it in flight records ev[9] and sets pinned_busy; the next user waits for
// that copy before writing into (or replacing) the buffer.
static int pinned(size_t bytes, void** out)
{
    if (g.pinned_busy) {
        HIPCHK(hipEventSynchronize(g.ev[9]));
        g.pinned_busy = false;
    }
    if (bytes > g.ws.pinned_cap) {
        if (g.ws.pinned) (void)hipHostFree(g.ws.pinned);
        g.ws.pinned = nullptr;
        g.ws.pinned_cap = 0;
        HIPCHK(hipHostMalloc(&g.ws.pinned, bytes, hipHostMallocDefault));
        g.ws.pinned_cap = bytes;
    }
    *out = g.ws.pinned;
    return D2BI_OK;
}

// ------------------------------------------------------------------ device atoms

}  // namespace d2bi

struct d2bi_atoms_s {
    double* x = nullptr;        // 3*n doubles, AoS
    int64_t n = 0;
    // unique per upload: a freed set's device address (and this struct's)
    // can come back for the next upload, with other positions
    uint64_t serial = 0;
};

namespace d2bi {

// ------------------------------------------------------------------ fills

// Fill kernels clear counters and flags (rather than hipMemsetAsync, which
// must not be recorded into a captured graph).
__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static void fill_i32(int* p, int64_t n, int v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// ------------------------------------------------------------------ generic scan

static constexpr int SCAN_BLOCK = 256;
static constexpr int SCAN_ITEMS = 4;
static constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Block-wide exclusive scan of one int per thread (wave64 shuffles + LDS).
__device__ inline int block_exclusive_scan(int v, int* total, int* lds /*[SCAN_BLOCK/64 + 1]*/)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) { int t = lds[w]; lds[w] = s; s += t; }
        lds[blockDim.x >> 6] = s;
    }
    __syncthreads();
    int r = incl - v + lds[wave];
    *total = lds[blockDim.x >> 6];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                            int* __restrict__ sums)
{
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) s += in[base + i];
    int tot;
    block_exclusive_scan(s, &tot, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single-block exclusive scan of the block sums (any length), in place.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int* __restrict__ sums, int64_t nb,
                                                          int* __restrict__ total_out)
{
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    int carry = 0;
    for (int64_t b = 0; b < nb; b += SCAN_BLOCK) {
        int64_t i = b + threadIdx.x;
        int v = i < nb ? sums[i] : 0;
        int tot;
        int e = block_exclusive_scan(v, &tot, lds);
        if (i < nb) sums[i] = e + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const int* __restrict__ in, int64_t n,
                                                           const int* __restrict__ sums,
                                                           int* __restrict__ out)
{
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int tot;
    int e = block_exclusive_scan(s, &tot, lds) + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) out[base + i] = e;
        e += v[i];
    }
}

static int64_t scan_tmp_ints(int64_t n)
{
    return std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE) + 1;
}

// out[0..n-1] = exclusive scan of in, out[n] = total.  in may equal out.
// sums: scan_tmp_ints(n) ints of scratch (no allocation here, so the call can
// be captured into a graph).
static int scan_exclusive_tmp(const int* in, int* out, int64_t n, int* sums, hipStream_t s)
{
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nb, out + n);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums, out);
    HIPCHK(hipGetLastError());
    return D2BI_OK;
}

static int scan_exclusive(const int* in, int* out, int64_t n, hipStream_t s, Workspace& w)
{
    CHK(w.scan_tmp.ensure(sizeof(int) * scan_tmp_ints(n)));
    return scan_exclusive_tmp(in, out, n, w.scan_tmp.as<int>(), s);
}

}  // namespace d2bi
