/*
 * d2bi_gpu.h -- C-ABI of the MI355X imaging library (libd2bi_gpu.so).
 *
 * The library replaces the hot path of d2bi's dynamical two-beam imaging:
 * atoms -> link-cell binning -> voxelised phase/density field ->
 * Howie-Whelan RK4 column integration -> image, plus the PNG blur.
 * Every entry point is plain C: pointers, sizes, structs of doubles and
 * 32/64-bit integers.  Nothing here exposes torch, HIP or Fortran
 * descriptor types, so a Fortran host binds it with iso_c_binding
 * (bind(c) interfaces mirror these declarations one for one).
 *
 * Conventions
 *   - 3x3 matrices are stored column-major (Fortran order): m[3*j+i] = M(i+1,j+1).
 *   - Images are (0:Mx-1, 0:My-1) column-major, x fastest, like the
 *     reference's d2big2 array.
 *   - Every function returns 0 on success and a negative D2BI_ERR_* code on
 *     failure; d2bi_last_error() gives a message.  Nothing calls exit/stop.
 *   - Host buffers belong to the caller; device buffers, streams and events
 *     belong to the library.  The library is not thread-safe: one host
 *     thread drives one device, as the reference's single-threaded rank does.
 */
#ifndef D2BI_GPU_H
#define D2BI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D2BI_MAX_UVW   125      /* Lib_DynamicalTwoBeamImaging.f90:86 uvw(3,125)          */
#define D2BI_MAX_PREC  256      /* precession points per condition                        */
#define D2BI_MAX_APERTURE 1024 /* aperture points per condition                         */
#define D2BI_SCRATCH_BYTES 67108864 /* 64 MiB: device workspace of d2bi_gpu_image_noalloc  */

enum {
    D2BI_OK            =  0,
    D2BI_ERR_ARG       = -1,    /* bad argument (null pointer, size, non-diagonal geometry) */
    D2BI_ERR_HIP       = -2,    /* a HIP runtime call failed                                */
    D2BI_ERR_CAPACITY  = -3,    /* a size limit was exceeded (see d2bi_last_error)          */
    D2BI_ERR_NOINIT    = -4,    /* d2bi_gpu_init was not called                             */
    D2BI_ERR_NODEVICE  = -5     /* no HIP device is visible                                 */
};

enum {
    D2BI_MODE_FP64 = 0,         /* reference arithmetic in FP64: membership bit-exact, 1e-10 */
    D2BI_MODE_FAST = 1          /* FP32 bulk, FP64 phase + FP64 guard-band membership        */
};

/*
 * The imaging parameter record: the fields of type DynamicalTwoBeamImaging
 * (Lib_DynamicalTwoBeamImaging.f90:64-90) that computeD2BI_img reads.  The
 * host fills it after its setup (setSupercellTilt*, setVoxelSupercellSize),
 * or with d2bi_setup_tilt_sg / d2bi_setup_tilt_ur below.
 */
typedef struct d2bi_record {
    double  a0;                 /* characteristic length; sigma = a0/2                 */
    double  a_cell[9];          /* conventional unit cell vectors (columns)            */
    double  g[3];               /* g-vector as reduced [hkl]                           */
    double  E;                  /* accelerating voltage (keV)                          */
    double  xi0, xig;           /* extinction distances (A)                            */
    double  U[9];               /* rotation taking zone axis to (0,0,1)                */
    double  R[9];               /* foil tilt                                           */
    double  delta[3];           /* offset after rotation                               */
    double  super_a[9];         /* input supercell: cell vectors ...                   */
    int32_t super_n[3];         /* ... and repeats (A_super = a * n)                   */
    int32_t nuvw;               /* number of periodic replicas                         */
    int32_t uvw[3 * D2BI_MAX_UVW];
    double  gk_a[9];            /* voxel cell vectors (diagonal: the cuboid / (Nx,Ny,Nz)) */
    int32_t gk_n[3];            /* voxel counts Mx, My, Mz                             */
    int32_t use_density;        /* 1: density field from nearest-atom range            */
    double  sg;                 /* deviation parameter at the untilted beam (1/A)      */
} d2bi_record;

/*
 * Derived quantities of the computeD2BI_img prologue
 * (Lib_DynamicalTwoBeamImaging.f90:1635-1836) for one serial patch.
 */
typedef struct d2bi_plan {
    int32_t Mx, My, Mz;         /* voxels                                              */
    int32_t deltaM;             /* voxels covering 3 sigma                             */
    int32_t nMax;               /* link-cell capacity estimate                         */
    int32_t Ex, Ey, Ez;         /* link cells                                          */
    int32_t nPrec;
    int32_t pad0;
    double  sigma, i2s2, rc2;   /* a0/2, 1/(2 sigma^2), (3 sigma)^2                    */
    double  a_col[9];           /* link-cell vectors                                   */
    double  ia_col[9];          /* their inverse (inverse3Mat)                         */
    double  lc_delta[3];        /* link-cell offset                                    */
    double  RU[9];              /* R*U                                                 */
    double  gg[3];              /* g in the lab frame (no beam tilt)                   */
    double  xmax[3];            /* image space (A)                                     */
    double  aperpx;             /* A per pixel                                         */
    double  modg;               /* |g| (1/A)                                           */
    double  dz, b0, bg;         /* slice thickness, pi/xi0, pi/xig                     */
    double  aa[D2BI_MAX_PREC];  /* 2 pi s_g for each precession point                  */
} d2bi_plan;

/* Per-call statistics, filled when a non-null pointer is passed. */
typedef struct d2bi_stats {
    int64_t n_points;           /* atoms in the link cells (after replicas/filter)     */
    int64_t n_guard;            /* FAST mode: candidate tests re-done in FP64          */
    int32_t max_per_cell;       /* largest link-cell population                        */
    int32_t n_cells;            /* Ex*Ey*Ez                                            */
    double  ms_bin;             /* device time of binning (ms)                         */
    double  ms_voxel;           /* device time of voxelisation (ms)                    */
    double  ms_integrate;       /* device time of column integration (ms)              */
    double  ms_total;           /* device time of the whole call (ms)                  */
    int64_t branch_bytes;       /* batches: device bytes of one branch's workspace     */
    int32_t n_branches;         /* batches: concurrent graph branches used             */
    int32_t pad0;
} d2bi_stats;

/* ---------------------------------------------------------------- runtime */

/* Select the HIP device and create the library's stream.  No reference
 * counterpart (the reference is CPU only); called once by the host. */
int         d2bi_gpu_init(int device);
int         d2bi_gpu_finalize(void);
const char* d2bi_last_error(void);
int         d2bi_version(void);

/* ---------------------------------------------------------------- host setup */

/* computeD2BI_img prologue (Lib_DynamicalTwoBeamImaging.f90:1635-1836),
 * serial patch geometry (factoriseParallel with one rank gives one patch,
 * Lib_FactoriseParallel.f90:104-106).  Pure host code. */
int d2bi_plan_image(const d2bi_record* rec, int64_t nAtoms,
                    double precAngle, int32_t nPrec, d2bi_plan* plan);

/* setSupercellTilt0 (Lib_DynamicalTwoBeamImaging.f90:567-788): find the tilt
 * about y giving the requested deviation parameter sg.  On entry rec holds
 * a0, a_cell, g, E, super_a/super_n; k is the zone axis [uvw]; mxy, mzz the
 * voxel subdivisions; d the voxel cuboid (all three > 0).  On return U, R,
 * delta, uvw, gk_a, gk_n and sg are set.  sg_target = NaN keeps zero tilt. */
int d2bi_setup_tilt_sg(d2bi_record* rec, const double k[3], double sg_target,
                       int32_t mxy, int32_t mzz, const double d[3]);

/* setSupercellTilt2 (Lib_DynamicalTwoBeamImaging.f90:1048-1106): use a
 * previously determined U and R (column-major). */
int d2bi_setup_tilt_ur(d2bi_record* rec, const double U[9], const double R[9],
                       int32_t mxy, int32_t mzz, const double d[3]);

/* setSupercellTilt1a without atom positions
 * (Lib_DynamicalTwoBeamImaging.f90:943-1045, with setSupercellTilt1 :795-939,
 * findBestRotationMatrix and tiltToDarkField,
 * Lib_DiffractionConditions.f90:129-357, 751-927): the (g, ng) two-beam
 * condition.  The crystal rotation U and tilt R put the reflection ng*g at
 * the Bragg condition with the weakest summed two-beam intensity of all other
 * permitted reflections within |hkl| <= 10, scored with the extinction
 * distances of the table xifile, then tilt about y to the dark-field minimum;
 * a fractional ng interpolates floor/ceiling by quaternion slerp.  On entry
 * rec holds a0, a_cell, g, E, xi0, xig (for g), super_a/super_n; lattice is
 * 1 sc, 2 bcc, 3 fcc, 4 hcp; d is the voxel cuboid.  On return U, R, delta,
 * uvw, gk_a, gk_n and sg are set and d[2] is the depth after the tilt
 * (voxelSupercellDimensions(3)).  ng >= 1. */
int d2bi_setup_tilt_ng(d2bi_record* rec, const double k[3], double ng, int32_t lattice,
                       const char* xifile, int32_t mxy, int32_t mzz, double d[3]);

/* rotateToStandardFrame (Lib_DiffractionConditions.f90:65-108). */
int d2bi_rotate_to_standard_frame(const double k[3], const double a_cell[9],
                                  const double hkl[3], double U[9]);

/* evenlySpacedPointsInCircle(x) (Lib_EvenlySpacedPointsInCircle.f90:29-100,
 * no file): n points in the unit disk, the reference's table of known
 * arrangements for n <= 25 (src/evenlySpaced.h, FP32 literals) and the
 * sunflower placement beyond.  xy is (2, n), x first. */
int d2bi_aperture_points(int32_t n, double* xy);

/* The aperture sampling of computeD2BI_img's optional apertureAngle and
 * nAperturePoints (Lib_DynamicalTwoBeamImaging.f90:1655-1667, 1805-1836,
 * findGdotR_aperture :1336-1344): gshift (3, nAperture) the g-vector used for
 * g.r at each aperture point, aa (nAperture, nPrec) 2 pi s_g of each
 * (aperture point, precession point), column-major. */
int d2bi_aperture_setup(const d2bi_record* rec, double precAngle, int32_t nPrec, double apertureAngle,
                        int32_t nAperture, double* gshift, double* aa);

/* wavelength / velocity (Lib_RelativisticElectrons.f90:24-50), V in volts. */
double d2bi_wavelength(double V);
double d2bi_velocity(double V);

/* getDeviationParameter(this,g) (Lib_DynamicalTwoBeamImaging.f90:268-282)
 * for a lab-frame g-vector. */
double d2bi_deviation_parameter(double E_keV, const double g_lab[3]);

/* ---------------------------------------------------------------- device atoms */

typedef struct d2bi_atoms_s* d2bi_atoms;

/* Copy atom positions x(1:3,1:nAtoms) (leading dimension ld >= 3, FP64,
 * column-major as in d2bi.f90:837-846) into HBM once, for reuse by many
 * conditions.  The reference keeps x on every rank after MPI_BCAST
 * (d2bi.f90:969-973). */
int d2bi_gpu_atoms_upload(const double* x, int64_t ld, int64_t nAtoms, d2bi_atoms* out);
int d2bi_gpu_atoms_free(d2bi_atoms atoms);
int64_t d2bi_gpu_atoms_count(d2bi_atoms atoms);

/* ---------------------------------------------------------------- imaging */

/* computeD2BI_img(this,x,d2big2,precessionAngle,nPrecessionAngle,...)
 * (Lib_DynamicalTwoBeamImaging.f90:1563), one serial patch, nAperturePoints=1
 * (the only value the d2bi command line can set, d2bi.f90:378-379).
 * img is Mx*My doubles (caller-owned), the precession average of |phi_g|^2.
 * x is host memory, copied to the device inside the call. */
int d2bi_gpu_image(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms,
                   double precAngle, int32_t nPrec, int32_t mode,
                   double* img, d2bi_stats* stats);

/* computeD2BI_img(this,x,d2big2,precessionAngle,nPrecessionAngle,
 *                 apertureAngle,nAperturePoints) (Lib_DynamicalTwoBeamImaging.f90:1563)
 * with the aperture sampled at nAperture points (radians; nAperture <=
 * D2BI_MAX_APERTURE): the mean over nAperture x nPrec integrations, each
 * aperture point with its own g.r (findGdotR_aperture :1269-1396).  The d2bi
 * command line keeps nAperture = 1 (d2bi.f90:92, 378-379); nAperture = 1
 * gives the image of d2bi_gpu_image. */
int d2bi_gpu_image_aperture(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms,
                            double precAngle, int32_t nPrec, double apertureAngle, int32_t nAperture,
                            int32_t mode, double* img, d2bi_stats* stats);

/* Same, with device-resident atoms: the timed region of a benchmark. */
int d2bi_gpu_image_dev(const d2bi_record* rec, d2bi_atoms atoms,
                       double precAngle, int32_t nPrec, int32_t mode,
                       double* img, d2bi_stats* stats);

/* Same, with the image left in HBM: d_img is a device pointer to Mx*My
 * doubles (e.g. a torch CUDA tensor), for GPU-side consumers (blur, further
 * conditions) and for timing the path without the PCIe copy of the image.
 * The call returns when the image is complete. */
int d2bi_gpu_image_dev_out(const d2bi_record* rec, d2bi_atoms atoms,
                           double precAngle, int32_t nPrec, int32_t mode,
                           double* d_img, d2bi_stats* stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc): the same
 * condition and image as d2bi_gpu_image_dev_out, bit for bit, with every
 * device buffer the call needs carved from a static block of
 * D2BI_SCRATCH_BYTES (64 MiB) per device, held from library load to process
 * end; the call itself calls no hipMalloc, hipFree or pinned-host allocator.
 * The caller passes the atoms (device-resident) and d_img (device memory,
 * Mx*My doubles).  A condition whose workspace exceeds the block is refused
 * with D2BI_ERR_CAPACITY (FAST needs 16 B and FP64 24 B per voxel for the
 * field, plus link cells and point records: e.g. 80 x 80 x 80 voxels with
 * 16,000 atoms fit).  Every call on a device shares that device's block and
 * overwrites it; calls that overlap on one device (from several host
 * threads) race.  Returns when the image is complete. */
int d2bi_gpu_image_noalloc(const d2bi_record* rec, d2bi_atoms atoms,
                           double precAngle, int32_t nPrec, int32_t mode,
                           double* d_img, d2bi_stats* stats);

/* Batched conditions: nRec records (tilts, g-vectors, s_g sweeps, ...)
 * over atom sets atoms[i] (entries may repeat), each integrated at nPrec[i]
 * precession points of angle precAngle[i].  img receives the images one
 * after another (sum of Mx*My over records).  The batch is executed as one
 * captured HIP graph per shape group; every record is re-voxelised, as the
 * reference re-voxelises each condition (d2bi.f90:1352-1394). */
int d2bi_gpu_image_batch(const d2bi_record* recs, const d2bi_atoms* atoms, int32_t nRec,
                         const double* precAngle, const int32_t* nPrec, int32_t mode,
                         double* img, d2bi_stats* stats);

/* The same batch (the d2bi driver's condition loop, d2bi.f90:1352-1394) with
 * the images left in HBM: d_img is a device pointer to the sum of Mx*My
 * doubles over the records, for GPU-side consumers and for timing the batch
 * without the PCIe copy of its images.  Returns when the images are
 * complete. */
int d2bi_gpu_image_batch_dev_out(const d2bi_record* recs, const d2bi_atoms* atoms, int32_t nRec,
                                 const double* precAngle, const int32_t* nPrec, int32_t mode,
                                 double* d_img, d2bi_stats* stats);

/* Voxel field for tests: per voxel (x fastest, then y, then z; Mz planes)
 * the normalised phase factor (re,im), density rho, neighbour count and the
 * minimum squared range.  Any output pointer may be null.  In FAST mode with
 * count and r2min both null the field comes from the imaging kernel itself
 * (otherwise from its diagnostic form, which also counts neighbours). */
int d2bi_gpu_field(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms,
                   int32_t mode, double* phase_re_im, double* rho,
                   int32_t* count, double* r2min);

/* ---------------------------------------------------------------- link cells */

/* findGdotR + LinkCell3D add (Lib_DynamicalTwoBeamImaging.f90:1160-1265,
 * Lib_LinkCell3D.f90:691-748): bin the rotated, replicated atoms of one
 * condition.  Outputs (any may be null): cell_start[Ex*Ey*Ez+1] (cell c =
 * cx + Ex*(cy + Ey*cz)), ids[cap] (1-based point number ii, insertion
 * order kept inside each cell), frac[3*cap] (offset from cell origin in cell
 * units), gdotr[cap] indexed by ii-1.  *nPoints returns Nr; cap is the size
 * of ids/frac/gdotr (error D2BI_ERR_CAPACITY if Nr > cap). */
int d2bi_gpu_linkcells(const d2bi_record* rec, const double* x, int64_t ld, int64_t nAtoms,
                       int32_t* cell_start, int32_t* ids, double* frac, double* gdotr,
                       int64_t cap, int64_t* nPoints);

/* Generic link cells: LinkCell3D_ctor0(a,nMax,Nx,Ny,Nz) + add1(x) +
 * neighbourList1b(x,rc) for each query point (Lib_LinkCell3D.f90:192,680,1515).
 * Points p(1:3,1:nP) get ids 1..nP.  For query q the neighbours are written
 * to id[q*maxNeigh ...], r2[...], with their number in n[q]. */
int d2bi_gpu_lc_query(const double a[9], int32_t Nx, int32_t Ny, int32_t Nz,
                      const double* p, int64_t nP, const double* q, int64_t nQ, double rc,
                      int32_t maxNeigh, int32_t* n, int32_t* id, double* r2);

/* ---------------------------------------------------------------- blur */

/* gaussianBlurImage(f_in,t,f_out) (d2bi.f90:1543-1624): separable,
 * edge-renormalised Gaussian of width t pixels, kernel half-width
 * max(5,ceiling(5t)). */
int d2bi_gpu_blur(const double* in, int32_t nx, int32_t ny, double t, double* out);

/* ---------------------------------------------------------------- synthetic input */

/* Seeded bcc crystal with thermal noise, Eshelby point dilatation centres
 * and spherical voids; the reference has no generator (inputs are files).
 * Writes up to cap atoms as x(1:3,:) (ld = 3) and returns the count in
 * *nAtoms; box[3] receives the periodic box lengths. */
typedef struct d2bi_synth_params {
    double   a0;                /* lattice parameter (A)                              */
    int32_t  nx, ny, nz;        /* cubic cells                                        */
    int32_t  n_centres;         /* Eshelby centres                                    */
    double   A_min, A_max;      /* dilatation strength range (A^3)                    */
    double   u_cap;             /* cap on |u| from one centre (A)                     */
    int32_t  n_voids;
    int32_t  pad0;
    double   r_min, r_max;      /* void radius range (A)                              */
    double   noise;             /* Gaussian sigma per coordinate (A)                  */
    double   offset[3];         /* added to every position                            */
    uint64_t seed;
} d2bi_synth_params;

int d2bi_synth_bcc(const d2bi_synth_params* p, double* x, int64_t cap,
                   int64_t* nAtoms, double box[3]);

#ifdef __cplusplus
}
#endif
#endif /* D2BI_GPU_H */
