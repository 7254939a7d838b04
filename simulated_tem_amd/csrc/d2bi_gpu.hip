// d2bi_gpu.hip -- gfx950 kernels and the device half of the C-ABI: one
// translation unit made of the parts included below (d2bi_*.cuh) and the
// C-ABI entry points at its end.
//
// Stages of one imaging condition (one "field"):
//   (a) binning      findGdotR + LinkCell3D add as a stable counting sort of
//                    the 27 x N replica predicate into link cells
//   (b) voxelise     per voxel S = sum_j exp(-r_j^2/2s^2 + i g.r_j) over the
//                    27 neighbour cells, rho from the nearest range
//   (c) integrate    RK4 Howie-Whelan march down each pixel column
//   (d) blur         separable Gaussian on the output image
// FP64 mode reproduces the reference arithmetic (un-fused, same order).
// FAST mode runs (b) and (c) in FP32 with FP64 phases and an FP64 re-test
// of every candidate whose FP32 range lies in a guard band around 3 sigma,
// so voxel membership stays identical to FP64.
//
// The unit is compiled with -ffp-contract=off: every fused multiply-add in it
// is written explicitly (fmaf) and only appears in FAST-mode FP32 code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

#include "../../include/d2bi_gpu.h"
#include "d2bi_internal.h"

// The parts, in order (one translation unit):
#include "d2bi_common.cuh"
#include "d2bi_bin.cuh"
#include "d2bi_axis.cuh"
#include "d2bi_fp64.cuh"
#include "d2bi_voxel_tile.cuh"
#include "d2bi_integrate_fast.cuh"
#include "d2bi_tune.cuh"
#include "d2bi_blur_query.cuh"
#include "d2bi_run.cuh"
#include "d2bi_batch.cuh"

// ------------------------------------------------------------------ C-ABI

using namespace d2bi;

// ------------------------------------------------------------------ no-allocation scratch

// The workspace of d2bi_gpu_image_noalloc: one static block in every device's
// copy of the code object, held from module load to process end.  Host code
// reaches a device's block only through the address hipGetSymbolAddress
// returns with that device current (g_scratch_dev), never through the
// symbol's name, which on the host is a stand-in and not a device address.
static __device__ __align__(256) char scratch[D2BI_SCRATCH_BYTES];
static constexpr int SCRATCH_MAX_DEVICES = 64;
static void* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The current device's scratch address, looked up once per device.
static int scratch_address(int device)
{
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        return seterr(D2BI_ERR_ARG, "device %d beyond the scratch table", device);
    if (g_scratch_dev[device]) return D2BI_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return seterr(D2BI_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch block");
    g_scratch_dev[device] = p;
    return D2BI_OK;
}

extern "C" {

int d2bi_version(void) { return 10000; }

const char* d2bi_last_error(void) { return g_err; }

int d2bi_gpu_init(int device)
{
    if (g.init) {
        if (device == g.device) return D2BI_OK;
        d2bi_gpu_finalize();
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return seterr(D2BI_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return seterr(D2BI_ERR_ARG, "device %d out of range (%d visible)", device, n);
    HIPCHK(hipSetDevice(device));
    // the no-allocation scratch's address on this device; should the lookup
    // fail, d2bi_gpu_image_noalloc repeats it and reports the error, and
    // every other entry point works as before
    if (scratch_address(device) != D2BI_OK) (void)hipGetLastError();
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    for (auto& e : g.ev) HIPCHK(hipEventCreate(&e));
    g.device = device;
    g.init = true;
    return D2BI_OK;
}

int d2bi_gpu_finalize(void)
{
    if (!g.init) return D2BI_OK;
    (void)hipSetDevice(g.device);
    (void)hipStreamSynchronize(g.stream);
    g_batch.release();
    g.ws.release_all();
    for (auto& e : g.ev) { (void)hipEventDestroy(e); e = nullptr; }
    (void)hipStreamDestroy(g.stream);
    g.stream = nullptr;
    g.init = false;
    g.device = -1;
    return D2BI_OK;
}

int d2bi_gpu_atoms_upload(const double* x, int64_t ld, int64_t nAtoms, d2bi_atoms* out)
{
    CHK(need_init());
    if (!out || nAtoms < 0 || (nAtoms > 0 && (!x || ld < 3))) return seterr(D2BI_ERR_ARG, "bad atom array");
    static uint64_t serial = 0;
    d2bi_atoms_s* a = new d2bi_atoms_s;
    a->n = nAtoms;
    a->serial = ++serial;
    size_t bytes = sizeof(double) * 3 * std::max<int64_t>(nAtoms, 1);
    if (hipMalloc(&a->x, bytes) != hipSuccess) {
        delete a;
        return seterr(D2BI_ERR_HIP, "hipMalloc for atoms failed");
    }
    std::vector<double> h(3 * std::max<int64_t>(nAtoms, 1));
    for (int64_t j = 0; j < nAtoms; j++) {
        h[3 * j] = x[ld * j]; h[3 * j + 1] = x[ld * j + 1]; h[3 * j + 2] = x[ld * j + 2];
    }
    if (hipMemcpy(a->x, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(a->x);
        delete a;
        return seterr(D2BI_ERR_HIP, "atom upload failed");
    }
    *out = a;
    return D2BI_OK;
}

int d2bi_gpu_atoms_free(d2bi_atoms atoms)
{
    if (!atoms) return D2BI_OK;
    if (atoms->x) (void)hipFree(atoms->x);
    delete atoms;
    return D2BI_OK;
}

int64_t d2bi_gpu_atoms_count(d2bi_atoms atoms) { return atoms ? atoms->n : -1; }

int d2bi_gpu_image_dev(const d2bi_record* rec, d2bi_atoms atoms, double precAngle, int32_t nPrec, int32_t mode,
                       double* img, d2bi_stats* stats)
{
    CHK(need_init());
    if (!rec || !atoms || !img) return seterr(D2BI_ERR_ARG, "null argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    d2bi_plan plan;
    CHK(image_condition(rec, atoms->x, atoms->n, precAngle, nPrec, mode, stats, &plan, g.ws));
    HIPCHK(hipMemcpy(img, g.ws.img.p, sizeof(double) * plan.Mx * plan.My, hipMemcpyDeviceToHost));
    return D2BI_OK;
}

int d2bi_gpu_image_dev_out(const d2bi_record* rec, d2bi_atoms atoms, double precAngle, int32_t nPrec,
                           int32_t mode, double* d_img, d2bi_stats* stats)
{
    CHK(need_init());
    if (!rec || !atoms || !d_img) return seterr(D2BI_ERR_ARG, "null argument");
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, d_img) != hipSuccess) {
        (void)hipGetLastError();                 // a failed query must not linger as the last error
        return seterr(D2BI_ERR_ARG, "d_img is not device memory");
    }
    if (attr.type != hipMemoryTypeDevice || attr.device != g.device)
        return seterr(D2BI_ERR_ARG, "d_img is not memory of device %d", g.device);
    if (stats) std::memset(stats, 0, sizeof *stats);
    d2bi_plan plan;
    CHK(image_condition(rec, atoms->x, atoms->n, precAngle, nPrec, mode, stats, &plan, g.ws, d_img));
    HIPCHK(hipStreamSynchronize(g.stream));
    return D2BI_OK;
}

int d2bi_gpu_image_noalloc(const d2bi_record* rec, d2bi_atoms atoms, double precAngle, int32_t nPrec,
                           int32_t mode, double* d_img, d2bi_stats* stats)
{
    CHK(need_init());
    if (!rec || !atoms || !d_img) return seterr(D2BI_ERR_ARG, "null argument");
    if (mode != D2BI_MODE_FP64 && mode != D2BI_MODE_FAST) return seterr(D2BI_ERR_ARG, "unknown mode %d", mode);
    CHK(scratch_address(g.device));                 // done at init unless it failed there
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, d_img) != hipSuccess) {
        (void)hipGetLastError();
        return seterr(D2BI_ERR_ARG, "d_img is not device memory");
    }
    if (attr.type != hipMemoryTypeDevice || attr.device != g.device)
        return seterr(D2BI_ERR_ARG, "d_img is not memory of device %d", g.device);
    // refuse, before anything is launched, a condition whose field, link
    // cells and replica masks alone exceed the scratch; the arena refuses
    // the rest (point arrays) when the point count is known
    d2bi_plan plan;
    CHK(plan_image(rec, atoms->n, precAngle, nPrec, &plan));
    const double nvox = (double)plan.Mx * plan.My * plan.Mz;
    const double ncell = (double)plan.Ex * plan.Ey * plan.Ez;
    const double need = nvox * (mode == D2BI_MODE_FP64 ? 24.0 : 16.0) + 8.0 * (ncell + 1) +
                        4.0 * (double)atoms->n;
    if (need > (double)D2BI_SCRATCH_BYTES)
        return seterr(D2BI_ERR_CAPACITY,
                      "condition of %d x %d x %d voxels needs at least %.0f bytes of workspace; the "
                      "no-allocation scratch holds %lld",
                      plan.Mx, plan.My, plan.Mz, need, (long long)D2BI_SCRATCH_BYTES);
    Arena ar;
    ar.base = static_cast<char*>(g_scratch_dev[g.device]);
    ar.size = (size_t)D2BI_SCRATCH_BYTES;
    Workspace w;
    w.bind(&ar);
    if (stats) std::memset(stats, 0, sizeof *stats);
    const int rc = image_condition(rec, atoms->x, atoms->n, precAngle, nPrec, mode, stats, &plan, w, d_img);
    // work already queued on the scratch finishes before the call returns,
    // also after an error
    const hipError_t e = hipStreamSynchronize(g.stream);
    if (rc != D2BI_OK) return rc;
    if (e != hipSuccess) return seterr(D2BI_ERR_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    return D2BI_OK;
}

int d2bi_gpu_image_aperture(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms,
                            double precAngle, int32_t nPrec, double apertureAngle, int32_t nAperture,
                            int32_t mode, double* img, d2bi_stats* stats)
{
    CHK(need_init());
    if (!rec || !img) return seterr(D2BI_ERR_ARG, "null argument");
    if (nAperture < 1 || nAperture > D2BI_MAX_APERTURE) return seterr(D2BI_ERR_ARG, "nAperture out of range");
    if (stats) std::memset(stats, 0, sizeof *stats);
    double* d_x;
    CHK(upload_atoms_tmp(x, ld, nAtoms, &d_x));
    d2bi_plan plan;
    CHK(image_aperture(rec, d_x, nAtoms, precAngle, nPrec, apertureAngle, nAperture, mode, stats, &plan));
    HIPCHK(hipMemcpy(img, g.ws.img.p, sizeof(double) * plan.Mx * plan.My, hipMemcpyDeviceToHost));
    return D2BI_OK;
}

int d2bi_gpu_image(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms, double precAngle,
                   int32_t nPrec, int32_t mode, double* img, d2bi_stats* stats)
{
    CHK(need_init());
    if (!rec || !img) return seterr(D2BI_ERR_ARG, "null argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    double* d_x;
    CHK(upload_atoms_tmp(x, ld, nAtoms, &d_x));
    d2bi_plan plan;
    CHK(image_condition(rec, d_x, nAtoms, precAngle, nPrec, mode, stats, &plan, g.ws));
    HIPCHK(hipMemcpy(img, g.ws.img.p, sizeof(double) * plan.Mx * plan.My, hipMemcpyDeviceToHost));
    return D2BI_OK;
}

int d2bi_gpu_image_batch(const d2bi_record* recs, const d2bi_atoms* atoms, int32_t nRec, const double* precAngle,
                         const int32_t* nPrec, int32_t mode, double* img, d2bi_stats* stats)
{
    CHK(need_init());
    if (nRec < 0 || (nRec > 0 && (!recs || !atoms || !precAngle || !nPrec || !img)))
        return seterr(D2BI_ERR_ARG, "null argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    for (int i = 0; i < nRec; i++)
        if (!atoms[i]) return seterr(D2BI_ERR_ARG, "null atom set %d", i);
    if (nRec == 0) return D2BI_OK;
    return run_batch(recs, atoms, nRec, precAngle, nPrec, mode, img, false, stats);
}

int d2bi_gpu_image_batch_dev_out(const d2bi_record* recs, const d2bi_atoms* atoms, int32_t nRec,
                                 const double* precAngle, const int32_t* nPrec, int32_t mode, double* d_img,
                                 d2bi_stats* stats)
{
    CHK(need_init());
    if (nRec < 0 || (nRec > 0 && (!recs || !atoms || !precAngle || !nPrec || !d_img)))
        return seterr(D2BI_ERR_ARG, "null argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    for (int i = 0; i < nRec; i++)
        if (!atoms[i]) return seterr(D2BI_ERR_ARG, "null atom set %d", i);
    if (nRec == 0) return D2BI_OK;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, d_img) != hipSuccess) {
        (void)hipGetLastError();                 // a failed query must not linger as the last error
        return seterr(D2BI_ERR_ARG, "d_img is not device memory");
    }
    if (attr.type != hipMemoryTypeDevice || attr.device != g.device)
        return seterr(D2BI_ERR_ARG, "d_img is not memory of device %d", g.device);
    return run_batch(recs, atoms, nRec, precAngle, nPrec, mode, d_img, true, stats);
}

int d2bi_gpu_field(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms, int32_t mode,
                   double* phase_re_im, double* rho, int32_t* count, double* r2min)
{
    CHK(need_init());
    if (!rec) return seterr(D2BI_ERR_ARG, "null argument");
    if (mode != D2BI_MODE_FP64 && mode != D2BI_MODE_FAST) return seterr(D2BI_ERR_ARG, "unknown mode %d", mode);
    double* d_x;
    CHK(upload_atoms_tmp(x, ld, nAtoms, &d_x));
    d2bi_plan plan;
    CHK(plan_image(rec, nAtoms, 0.0, 1, &plan));
    AxisTables tab;
    CHK(axis_tables(rec, &plan, &tab));
    BinParams BP;
    CHK(bin_params(rec, &plan, &BP));
    BP.fast = mode == D2BI_MODE_FAST ? 1 : 0;
    hipStream_t s = g.stream;
    AxisDev ax;
    CHK(upload_axis(tab, &ax, s, g.ws));
    Binned b;
    CHK(run_binning(d_x, nAtoms, BP, &b, s, g.ws));
    // FAST without count and r2min runs the imaging kernel itself (the
    // variant imaging uses)
    const bool diag = mode == D2BI_MODE_FP64 || count || r2min;
    CHK(run_voxel(rec, &plan, tab, ax, mode, diag, nullptr, s, g.ws));
    HIPCHK(hipStreamSynchronize(s));
    const int64_t nvox = (int64_t)plan.Mx * plan.My * plan.Mz;
    Workspace& w = g.ws;
    if (mode == D2BI_MODE_FP64) {
        if (phase_re_im) {
            std::vector<double> re(nvox), im(nvox);
            HIPCHK(hipMemcpy(re.data(), w.f64_re.p, sizeof(double) * nvox, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(im.data(), w.f64_im.p, sizeof(double) * nvox, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < nvox; i++) { phase_re_im[2 * i] = re[i]; phase_re_im[2 * i + 1] = im[i]; }
        }
        if (rho) HIPCHK(hipMemcpy(rho, w.f64_rho.p, sizeof(double) * nvox, hipMemcpyDeviceToHost));
        if (r2min) HIPCHK(hipMemcpy(r2min, w.r2m.p, sizeof(double) * nvox, hipMemcpyDeviceToHost));
    } else {
        // (re, im, rho, 0) per voxel, or (re, im) from the imaging kernel
        // without the density field (rho = 1)
        const int nc = (!diag && !rec->use_density) ? 2 : 4;
        std::vector<float> f(nc * nvox);
        HIPCHK(hipMemcpy(f.data(), w.f32_field.p, sizeof(float) * nc * nvox, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nvox; i++) {
            if (phase_re_im) { phase_re_im[2 * i] = f[nc * i]; phase_re_im[2 * i + 1] = f[nc * i + 1]; }
            if (rho) rho[i] = nc == 4 ? f[4 * i + 2] : 1.0f;
        }
        if (r2min) {
            std::vector<float> r(nvox);
            HIPCHK(hipMemcpy(r.data(), w.r2m.p, sizeof(float) * nvox, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < nvox; i++) r2min[i] = r[i];
        }
    }
    if (count) HIPCHK(hipMemcpy(count, w.cnt.p, sizeof(int) * nvox, hipMemcpyDeviceToHost));
    return D2BI_OK;
}

int d2bi_gpu_linkcells(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms, int32_t* cell_start,
                       int32_t* ids, double* frac, double* gdotr, int64_t cap, int64_t* nPoints)
{
    CHK(need_init());
    if (!rec) return seterr(D2BI_ERR_ARG, "null argument");
    double* d_x;
    CHK(upload_atoms_tmp(x, ld, nAtoms, &d_x));
    d2bi_plan plan;
    CHK(plan_image(rec, nAtoms, 0.0, 1, &plan));
    BinParams BP;
    CHK(bin_params(rec, &plan, &BP));
    Binned b;
    CHK(run_binning(d_x, nAtoms, BP, &b, g.stream, g.ws));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (nPoints) *nPoints = b.nr;
    if ((ids || frac || gdotr) && b.nr > cap)
        return seterr(D2BI_ERR_CAPACITY, "%lld points exceed cap %lld", (long long)b.nr, (long long)cap);
    Workspace& w = g.ws;
    if (cell_start) HIPCHK(hipMemcpy(cell_start, w.cell_start.p, sizeof(int) * (b.ncell + 1), hipMemcpyDeviceToHost));
    if (ids && b.nr) HIPCHK(hipMemcpy(ids, w.slot_id.p, sizeof(int) * b.nr, hipMemcpyDeviceToHost));
    if (frac && b.nr) {
        std::vector<double> f(4 * b.nr);
        HIPCHK(hipMemcpy(f.data(), w.slot_frac.p, sizeof(double) * 4 * b.nr, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < b.nr; i++) { frac[3 * i] = f[4 * i]; frac[3 * i + 1] = f[4 * i + 1]; frac[3 * i + 2] = f[4 * i + 2]; }
    }
    if (gdotr && b.nr) HIPCHK(hipMemcpy(gdotr, w.pt_gdotr.p, sizeof(double) * b.nr, hipMemcpyDeviceToHost));
    return D2BI_OK;
}

int d2bi_gpu_lc_query(const double a[9], int32_t Nx, int32_t Ny, int32_t Nz, const double* p, int64_t nP,
                      const double* q, int64_t nQ, double rc, int32_t maxNeigh, int32_t* n, int32_t* id, double* r2)
{
    CHK(need_init());
    if (!a || Nx < 1 || Ny < 1 || Nz < 1 || nP < 0 || nQ < 0 || maxNeigh < 0 || (nQ > 0 && (!q || !n)) ||
        (nP > 0 && !p) || (maxNeigh > 0 && nQ > 0 && (!id || !r2)))
        return seterr(D2BI_ERR_ARG, "bad argument");
    hipStream_t s = g.stream;
    double* d_x;
    CHK(upload_atoms_tmp(p, 3, nP, &d_x));
    BinParams BP;
    std::memset(&BP, 0, sizeof BP);
    double ia[9];
    inverse3(a, ia);
    std::memcpy(BP.ia, ia, sizeof ia);
    BP.nuvw = 1;
    BP.Ex = Nx; BP.Ey = Ny; BP.Ez = Nz;
    BP.generic = 1;
    BP.cell_a[0] = a[0]; BP.cell_a[1] = a[4]; BP.cell_a[2] = a[8];
    Binned b;
    CHK(run_binning(d_x, nP, BP, &b, s, g.ws));
    QueryParams Q;
    std::memcpy(Q.a, a, sizeof Q.a);
    std::memcpy(Q.ia, ia, sizeof Q.ia);
    Q.Nx = Nx; Q.Ny = Ny; Q.Nz = Nz;
    Q.rc2 = rc * rc;
    Q.maxNeigh = maxNeigh;
    Workspace& w = g.ws;
    if (nQ > 0) {
        Buf dq, dn, did, dr2;
        int rc_ = D2BI_OK;
        if ((rc_ = dq.ensure(sizeof(double) * 3 * nQ)) || (rc_ = dn.ensure(sizeof(int) * nQ)) ||
            (rc_ = did.ensure(sizeof(int) * std::max<int64_t>(1, nQ * maxNeigh))) ||
            (rc_ = dr2.ensure(sizeof(double) * std::max<int64_t>(1, nQ * maxNeigh)))) {
            dq.release(); dn.release(); did.release(); dr2.release();
            return rc_;
        }
        hipError_t e = hipMemcpy(dq.p, q, sizeof(double) * 3 * nQ, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_lc_query, dim3((unsigned)((nQ + 255) / 256)), dim3(256), 0, s, Q, nQ,
                               dq.as<double>(), w.cell_start.as<int>(), w.slot_id.as<int>(),
                               w.slot_frac.as<double4>(), dn.as<int>(), did.as<int>(), dr2.as<double>());
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipMemcpy(n, dn.p, sizeof(int) * nQ, hipMemcpyDeviceToHost);
        if (e == hipSuccess && maxNeigh > 0) e = hipMemcpy(id, did.p, sizeof(int) * nQ * maxNeigh, hipMemcpyDeviceToHost);
        if (e == hipSuccess && maxNeigh > 0) e = hipMemcpy(r2, dr2.p, sizeof(double) * nQ * maxNeigh, hipMemcpyDeviceToHost);
        dq.release(); dn.release(); did.release(); dr2.release();
        if (e != hipSuccess) return seterr(D2BI_ERR_HIP, "link-cell query failed: %s", hipGetErrorString(e));
    }
    return D2BI_OK;
}

int d2bi_gpu_blur(const double* in, int32_t nx, int32_t ny, double t, double* out)
{
    CHK(need_init());
    if (!in || !out || nx < 1 || ny < 1 || !(t > 0)) return seterr(D2BI_ERR_ARG, "bad blur argument");
    int nk = (int)std::ceil(t * 5);
    if (nk < 5) nk = 5;
    if (nk > BLUR_MAXK) return seterr(D2BI_ERR_CAPACITY, "blur half-width %d exceeds %d", nk, BLUR_MAXK);
    std::vector<double> kern(nk + 1);
    const double i2s2 = 1 / (2 * t * t);
    kern[0] = 1.0;
    for (int i = 1; i <= nk; i++) kern[i] = std::exp((double)(-i * i) * i2s2);
    // edge weight sums W(i) = sum of kern[|j - i|] over the window inside the
    // image, j ascending (d2bi.f90:1585-1590, 1609-1613): 1 / W_x per column,
    // W_y per row
    auto wsum = [&](int i, int n) {
        double ws = 0.0;
        for (int j = std::max(0, i - nk); j <= std::min(n - 1, i + nk); j++) ws = ws + kern[std::abs(j - i)];
        return ws;
    };
    std::vector<double> wx_inv(nx), wy(ny);
    for (int i = 0; i < nx; i++) wx_inv[i] = 1 / wsum(i, nx);
    for (int i = 0; i < ny; i++) wy[i] = wsum(i, ny);
    Workspace& w = g.ws;
    const int64_t np = (int64_t)nx * ny;
    CHK(w.blur_in.ensure(sizeof(double) * np));
    CHK(w.blur_tmp.ensure(sizeof(double) * np));
    CHK(w.blur_out.ensure(sizeof(double) * np));
    CHK(w.blur_w.ensure(sizeof(double) * nx));
    CHK(w.blur_wy.ensure(sizeof(double) * ny));
    CHK(w.blur_k.ensure(sizeof(double) * (nk + 1)));
    hipStream_t s = g.stream;
    HIPCHK(hipMemcpyAsync(w.blur_in.p, in, sizeof(double) * np, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.blur_k.p, kern.data(), sizeof(double) * (nk + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.blur_w.p, wx_inv.data(), sizeof(double) * nx, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.blur_wy.p, wy.data(), sizeof(double) * ny, hipMemcpyHostToDevice, s));
    BlurParams B{nx, ny, nk};
    hipLaunchKernelGGL(k_blur_x, dim3((nx + BLUR_TILE - 1) / BLUR_TILE, ny), dim3(BLUR_TILE), 0, s, B,
                       w.blur_in.as<double>(), w.blur_k.as<double>(), w.blur_tmp.as<double>());
    hipLaunchKernelGGL(k_blur_y, dim3((nx + 63) / 64, (ny + 4 * BLUR_RY - 1) / (4 * BLUR_RY)), dim3(256), 0, s, B,
                       w.blur_tmp.as<double>(), w.blur_k.as<double>(), w.blur_w.as<double>(), w.blur_wy.as<double>(),
                       w.blur_out.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, w.blur_out.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return D2BI_OK;
}

}  // extern "C"
