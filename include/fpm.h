/*
 * fpm.h - C ABI of the MI355X particle-mesh library (libfpm.so).
 *
 * Every compute entry point enqueues work on the given hipStream_t and returns
 * an int status (FPM_OK or a negative FPM_ERR_* code); none synchronises the
 * stream and none allocates device memory.  Device pointers are owned by the
 * caller.  Each entry point exists for float (_f32) and double (_f64) storage;
 * the pointer arguments typed void* hold that real type (or its complex pair).
 *
 * Mesh layouts
 *   real field     : [n0][n1][n2] C order, unpadded.
 *   complex field  : [n0][n1][pitch2] interleaved (re, im) pairs,
 *                    pitch2 >= n2/2 + 1; mode (i, j, k) has wave numbers
 *                    2 pi fftfreq(i)/L0, 2 pi fftfreq(j)/L1, 2 pi k/L2.
 *   The forward transform is normalised by 1/(n0 n1 n2), the inverse is not.
 *
 * Particle layout
 *   Positions are X = Q + S with S a [n][3] displacement array and Q either an
 *   explicit [n][3] array or, when q == NULL, the implicit lattice
 *   Q_p = (idx(p) + shift) * BoxSize / lattice, idx(p) the C-order index of p
 *   on a lattice[0] x lattice[1] x lattice[2] grid.
 *
 * Each declaration cites the reference operation it replaces
 * (rainwoodman/fastpm-python; pmesh is the reference's mesh dependency).
 */
#ifndef FPM_H
#define FPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FPM_OK 0
#define FPM_ERR_ARG (-1)          /* bad shape, size, order or pointer */
#define FPM_ERR_UNSUPPORTED (-2)  /* FFT length without a compiled plan */
#define FPM_ERR_HIP (-3)          /* a HIP launch failed */
#define FPM_ERR_WORKSPACE (-4)    /* workspace too small */
#define FPM_ERR_SCRATCH (-5)      /* beyond the fixed scratch of an allocation-free call */

typedef struct hipStream_t_opaque* fpm_stream_t; /* ABI-compatible with hipStream_t */

typedef struct fpm_particles {
  const void* q;          /* explicit Lagrangian positions [n][3], or NULL for the lattice */
  const void* s;          /* displacement [n][3], or NULL for zero; positions are
                             wrapped periodically and must lie within 2^31 cells
                             of the box */
  int64_t n;              /* particle count (0 is accepted everywhere: paints
                             write a zero mesh, particle outputs may be NULL) */
  int32_t lattice[3];     /* lattice particles per side (used when q == NULL) */
  int32_t pad_;
  double lattice_shift;   /* lattice offset in lattice spacings */
} fpm_particles;

typedef struct fpm_mesh {
  int32_t nmesh[3];
  int32_t variant;        /* kernel variants for this mesh, FPM_VAR_* fields; 0 = defaults */
  double box[3];
} fpm_mesh;

/* Fields of fpm_mesh.variant (2 bits each).  Every choice computes the same
 * results (bitwise for the particle kernels; to rounding for the transform
 * tilings, whose per-column arithmetic is the same up to multiply-add
 * contraction); fastpm.tune times them per mesh and keeps the fastest.  No
 * counterpart in the reference (pmesh plans its own transforms).
 *   FPM_VAR_FFT_COLS : strided transform passes, 0 = the process-wide choice
 *                      of fpm_fft_set_tiling, 1 = 64-byte, 2 = 128-byte row
 *                      segments per column tile
 *   FPM_VAR_Z_LINES  : z transform pass, 0 = 2048, 1 = 1024, 2 = 4096 complex
 *                      elements (rows x length) per workgroup
 *   FPM_VAR_BIN      : binning walk over lattice particles, 0 = automatic
 *                      (4 x 8 x 16 blocks where the lattice divides, else
 *                      8 x 8 patches), 1 = patches, 2 = index order
 *   FPM_VAR_READOUT  : threads per tile of the stencil force readout
 *                      (fpm_readout_tiled_grad4_f32), 0 = default (1024 CIC,
 *                      512 TSC), 1 = 256, 2 = 512, 3 = 1024
 *   FPM_VAR_PAINT    : threads per tile of the tiled paint, 0 = default
 *                      (512), 1 = 256, 2 = 1024
 *   FPM_VAR_RHO      : the stencil readout's RHO, 0 = from a density box
 *                      staged in LDS, 1 = read from the mesh through the
 *                      cache (a third less LDS per tile)
 *   FPM_VAR_SPAN     : tiles per workgroup of the stencil readout, 0 = one,
 *                      1 = two adjacent along z sharing one pair of boxes
 *                      (where the tile count along z is even; at least 512
 *                      threads) */
#define FPM_VAR_FFT_COLS 0
#define FPM_VAR_Z_LINES 2
#define FPM_VAR_BIN 4
#define FPM_VAR_READOUT 6
#define FPM_VAR_PAINT 8
#define FPM_VAR_RHO 10
#define FPM_VAR_SPAN 12
#define FPM_VAR_FIELD(variant, shift) (((variant) >> (shift)) & 3)

/* A k-space multiplier, evaluated per mode and fused into the first pass of the
 * inverse transform:  factor * [1/k^2 if laplace, 0 at k=0]
 *                     * prod_g (i a_{dir_g}(k))  (g < ngrad)
 *                     * exp(-k^2 r_split^2)       (if r_split != 0)
 *                     / prod_d W(k_d)^deconv      (window compensation)
 * with a_d(k) = k_d (grad_order 0, zero on the Nyquist plane) or the 4-point
 * finite difference (8 sin(k h) - sin(2 k h)) / (6 h), h = L/N (grad_order 1),
 * and W(k) = sinc(k h / 2).
 * Reference: fastpm/force/kernels.py:3-35 (laplace, gradient, longrange). */
typedef struct fpm_kop {
  double factor;
  double r_split;
  int32_t laplace;
  int32_t ngrad;
  int32_t grad_dir[2];
  int32_t grad_order;
  int32_t deconv;
} fpm_kop;

const char* fpm_version(void);
const char* fpm_strerror(int status);

/* Twiddle table for a transform axis of length n: n entries exp(-2 pi i m / n),
 * computed in double on the host and rounded to the storage type.  Writes to
 * host memory; the caller uploads it.  Returns FPM_ERR_UNSUPPORTED for a
 * length without a compiled plan. */
int fpm_fft_supported(int n);
int fpm_twiddles_f32(int n, float* host_out);
int fpm_twiddles_f64(int n, double* host_out);

/* Process-wide default tiling of the strided (y, x) transform passes, used
 * by meshes whose FPM_VAR_FFT_COLS field is 0: for axis length n and element
 * size elem_bytes (4: float32, 8: float64), wide = 1 selects column tiles of
 * twice the default row segment (128 instead of 64 bytes), wide = 0 the
 * default.  Results agree to rounding (each column is transformed with the
 * same arithmetic, up to multiply-add contraction).  Set it before launching
 * transforms from other threads; a mesh's own variant field overrides it.
 * No counterpart in the reference (pmesh plans its FFTs through pfft). */
int fpm_fft_set_tiling(int elem_bytes, int n, int wide);
int fpm_fft_get_tiling(int elem_bytes, int n);

/* Mass assignment.  Replaces pmesh ParticleMesh.paint / RealField.paint
 * (fastpm/core.py:191 prepare_force, fastpm/core.py:54 to_mesh).
 * order 2 = CIC, 3 = TSC.  Each particle deposits `mass` with periodic wrap.
 * accumulate = 0 clears `out` first (hold=False); 1 adds to it (hold=True). */
int fpm_paint_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                  void* out, int accumulate, fpm_stream_t stream);
int fpm_paint_f64(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                  void* out, int accumulate, fpm_stream_t stream);

/* Window interpolation of ncomp (<= 4) meshes at the particle positions, fused
 * into one gather.  Replaces RealField.readout (fastpm/core.py:203,
 * fastpm/force/gravity.py:24, fastpm/force/lpt.py:21).
 * outs[c][p * out_strides[c]] = scales[c] * sum_w W * meshes[c][cell]. */
int fpm_readout_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                    const void* const* meshes, void* const* outs, const int64_t* out_strides,
                    const double* scales, fpm_stream_t stream);
int fpm_readout_f64(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                    const void* const* meshes, void* const* outs, const int64_t* out_strides,
                    const double* scales, fpm_stream_t stream);

/* Tile-binned paint and readout (float32 storage, meshes whose sides are
 * multiples of 16).  The mesh is cut into 16^3-cell tiles; fpm_bin_particles
 * sorts the particles into per-tile lists held in the caller's workspace
 * (fpm_tile_workspace_size bytes): each slot holds the particle index and a
 * record of its window (offsets within the first cell, tile-local origin).
 * fpm_paint_tiled and fpm_readout_tiled read the lists only -- the positions
 * as of the bin call are used -- so they must follow a bin call with the same
 * particles, mesh and order.  The tiled paint accumulates in 64-bit fixed
 * point (each deposit rounded to a step of 2^-30 of |mass| or finer), is
 * bitwise reproducible, overwrites `out` (hold=False) and needs no
 * clearing pass.  Same contracts as fpm_paint / fpm_readout
 * (fastpm/core.py:191,203; fastpm/force/gravity.py:24).
 * Every bin call also leaves in the workspace a capacity layout for the
 * next binning of as many particles: each sub-bin (tile and crossing mask)
 * gets count + count/4 + 32 slots of this call's count (count + count/2 + 64
 * for the four layouts after a call that had to place its records exactly).
 * fpm_rebin_particles
 * places the records in one pass into that layout, straight at their
 * sub-bin's start plus their rank (no count/scatter pair); it needs the
 * workspace untouched since the bin or rebin call that left the layout.
 * When a sub-bin outgrows its capacity, or the layout is for another
 * particle count or mesh, it places every record exactly as
 * fpm_bin_particles does.  The paints and readouts give the same results
 * after either call.  No counterpart in the reference (pmesh paints with
 * atomics and keeps no lists).
 * fpm_tile_misplaced_offset gives the byte offset in the workspace of an int32
 * counter, zeroed by each bin call, of particles the scatter could not store
 * (a slot outside [0, n), which only a disagreement between the count and
 * scatter passes could produce); nonzero means the lists are incomplete and
 * the tiled kernels must not be used with them.  0 is the only valid
 * outcome.  The int32 after it is nonzero when the last bin or rebin call
 * placed the records exactly, zero when it placed them in one pass. */
int fpm_tiles_supported(const fpm_mesh* mesh);
int64_t fpm_tile_workspace_size(const fpm_mesh* mesh, int order, int64_t n);
int64_t fpm_tile_misplaced_offset(const fpm_mesh* mesh, int order, int64_t n);
int fpm_bin_particles_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* ws,
                          int64_t ws_bytes, fpm_stream_t stream);
int fpm_rebin_particles_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* ws,
                            int64_t ws_bytes, fpm_stream_t stream);

/* A leapfrog step boundary fused into the next Force's binning: equal to
 * fpm_kick_drift_f32(part->s, P, F, 3 n, kicks, nkick, drift) followed by
 * fpm_rebin_particles_f32(part, ...), bit for bit (the same updates of P and
 * of the displacement S = part->s, which must not be NULL and is written;
 * the same binning of the new positions), in one pass over the particles
 * that bins each particle straight from its new displacement instead of
 * reading it back.  Replaces the Kick, [Kick,] Drift of
 * fastpm/core.py:168-180 and the binning of the Force's paint
 * (fastpm/core.py:191) that follows them. */
int fpm_kick_drift_rebin_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* P,
                             const void* F, const double* kicks, int nkick, double drift,
                             void* ws, int64_t ws_bytes, fpm_stream_t stream);
int fpm_paint_tiled_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                        void* out, void* ws, int64_t ws_bytes, fpm_stream_t stream);
int fpm_readout_tiled_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                          const void* const* meshes, void* const* outs, const int64_t* out_strides,
                          const double* scales, void* ws, int64_t ws_bytes, fpm_stream_t stream);

/* Weighted tiled paint of ncomp (<= 3) weight arrays at once:
 * outs[c](cell) = scale * sum_p weights[c][p * wstrides[c]] W(x_p - cell), the
 * adjoint of a readout (readout_gradient of fastpm/operators.py:236-247, as
 * used by gravity_gradient).  Uses the binning of the last
 * fpm_bin_particles_f32 call; accumulates each component in 64-bit fixed
 * point in LDS, scaled per tile by the count of depositing particles and
 * their largest |weight|, so the result is bitwise reproducible; weights
 * must be finite; each output cell is written once. */
int fpm_paint_tiled_weighted_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                 int ncomp, const void* const* weights, const int64_t* wstrides,
                                 double scale, void* const* outs, void* ws, int64_t ws_bytes,
                                 fpm_stream_t stream);

/* Force readout with the four-point gradient applied in real space
 * (replaces fastpm/force/gravity.py:17-24 -- c2r(i a_d(k) pot_k) and a readout
 * for each d -- and the RHO readout of fastpm/core.py:203).  The four-point
 * gradient a_d = (8 sin k_d h - sin 2 k_d h) / (6 h) is the Fourier symbol of
 * the stencil [8 (f(x+h) - f(x-h)) - (f(x+2h) - f(x-2h))] / (12 h), so with
 * phi = c2r(pot_k) this writes
 *   force_out[p * force_stride + d] = force_scale * readout(D_d phi)(x_p)
 * and, when rho_out is not NULL, rho_out[p * rho_stride] = rho_scale *
 * readout(rho)(x_p).  Same binning workspace as fpm_readout_tiled_f32; float32
 * meshes that are multiples of 16 and at least 19 (CIC) / 20 (TSC) cells per
 * side. */
int fpm_readout_tiled_grad4_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                const void* rho, const void* phi, void* rho_out,
                                int64_t rho_stride, void* force_out, int64_t force_stride,
                                double rho_scale, double force_scale, void* ws, int64_t ws_bytes,
                                fpm_stream_t stream);

/* Position gradient of the stencil force readout: with D_d the four-point
 * stencil of fpm_readout_tiled_grad4_f32 (including 1 / (12 h_d)),
 *   gx[3 p + e] (+)= scale * d/dx_e sum_d gf[p * gf_stride + d] readout(D_d phi)(x_p)
 * (accumulate selects += or =), i.e. fpm_window_grad of the three meshes
 * D_d phi weighted by gf, without forming those meshes.  The window-derivative
 * part of gravity_gradient (fastpm/operators.py:236-247) for the stencil force.
 * Uses the binning of the last fpm_bin_particles_f32 call; same mesh
 * requirements as fpm_readout_tiled_grad4_f32. */
int fpm_window_grad_tiled_grad4_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                    const void* phi, const void* gf, int64_t gf_stride, void* gx,
                                    double scale, int accumulate, void* ws, int64_t ws_bytes,
                                    fpm_stream_t stream);

/* Real-to-half-complex 3D transform, normalised by 1/(n0 n1 n2).  `in` is left
 * intact.  tw0/tw1/tw2: device twiddle tables for n0, n1, n2.
 * Replaces RealField.r2c (fastpm/core.py:205, fastpm/force/lpt.py:64). */
int fpm_fft_r2c_f32(const void* in, void* out, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);
int fpm_fft_r2c_f64(const void* in, void* out, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);

/* Multiplier-then-inverse transform for ncomp (<= 6) outputs from one input:
 * outs[c] = c2r(ops[c](k) * in).  The input is read once; work[c] is a
 * complex scratch mesh of the same shape per output.  `in` is left intact.
 * ops == NULL means the identity multiplier.
 * Replaces ComplexField.apply(...).c2r (fastpm/force/gravity.py:17-22,
 * fastpm/force/lpt.py:18-20,39-57). */
int fpm_fft_c2r_f32(const void* in, int ncomp, const fpm_kop* ops, void* const* outs,
                    void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);
int fpm_fft_c2r_f64(const void* in, int ncomp, const fpm_kop* ops, void* const* outs,
                    void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);

/* Forward transform followed by multiplier-then-inverse transforms:
 * kout = r2c(in) (normalised by 1/(n0 n1 n2), as fpm_fft_r2c) and
 * outs[c] = c2r(ops[c](k) * kout) for ncomp (<= 6) outputs, with work[c]
 * complex scratch meshes as in fpm_fft_c2r.  The last forward pass and the
 * first inverse pass run along the same axis, so one kernel does both and
 * the spectrum is read once less than with separate calls.  `in` is left
 * intact (no work[c] may alias it).  kout may be NULL when the spectrum
 * itself is not wanted: it is then not stored (one half-spectrum write
 * less).  The forward passes run in work[0] (in kout instead where n2/2 or
 * n1 is not a multiple of 16); kout, when given, receives the spectrum in
 * the usual layout.  Replaces the Force's rho.r2c (fastpm/core.py:205) followed by
 * pot_k = delta_k.apply(laplace).apply(longrange) and its c2r
 * (fastpm/force/gravity.py:14-22). */
int fpm_fft_r2c_c2r_f32(const void* in, void* kout, int ncomp, const fpm_kop* ops,
                        void* const* outs, void* const* work, const fpm_mesh* mesh,
                        int64_t pitch2, const void* tw0, const void* tw1, const void* tw2,
                        fpm_stream_t stream);
int fpm_fft_r2c_c2r_f64(const void* in, void* kout, int ncomp, const fpm_kop* ops,
                        void* const* outs, void* const* work, const fpm_mesh* mesh,
                        int64_t pitch2, const void* tw0, const void* tw1, const void* tw2,
                        fpm_stream_t stream);

/* The three gradient components of one field in one call:
 * outs[d] = c2r(i a_d(k) * common(k) * in), d = 0, 1, 2, where `common` has
 * ngrad == 0 and its grad_order selects a_d.  Uses that a_d depends on k_d
 * only, so the shared product is transformed once along x and y: eight mesh
 * passes instead of ten.  work[0..2]: distinct complex scratch meshes.
 * Replaces the per-axis loop of fastpm/force/gravity.py:17-24. */
int fpm_fft_c2r_grad3_f32(const void* in, const fpm_kop* common, void* const* outs,
                          void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                          const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);
int fpm_fft_c2r_grad3_f64(const void* in, const fpm_kop* common, void* const* outs,
                          void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                          const void* tw0, const void* tw1, const void* tw2, fpm_stream_t stream);

/* In-place k-space multiplier on a complex field (ComplexField.apply with one
 * of the kernels of fastpm/force/kernels.py:3-35). */
int fpm_kspace_apply_f32(void* field, const fpm_kop* op, const fpm_mesh* mesh, int64_t pitch2,
                         fpm_stream_t stream);
int fpm_kspace_apply_f64(void* field, const fpm_kop* op, const fpm_mesh* mesh, int64_t pitch2,
                         fpm_stream_t stream);

/* Complex white noise with <|w|^2> = 1 (|w| = 1 if unitary), Hermitian
 * consistent on the k_z = 0 and Nyquist planes, zero mode 0.  Counter-based
 * (Philox4x32-10 keyed by seed, counter = canonical mode index), so the field
 * does not depend on launch shape.
 * Replaces pmesh generate_whitenoise(seed, type='complex') (fastpm/core.py:114). */
int fpm_whitenoise_f32(void* out, const fpm_mesh* mesh, int64_t pitch2, uint64_t seed,
                       int unitary, fpm_stream_t stream);
int fpm_whitenoise_f64(void* out, const fpm_mesh* mesh, int64_t pitch2, uint64_t seed,
                       int unitary, fpm_stream_t stream);

/* delta_k *= table[nx^2 + ny^2 + nz^2] (integer wave numbers), with the device
 * table holding sqrt(P(k)/V) at k = 2 pi sqrt(n^2)/L for an isotropic box.
 * Replaces Solver.linear (fastpm/core.py:116-118). */
int fpm_linear_scale_f32(void* field, const fpm_mesh* mesh, int64_t pitch2, const double* table,
                         int64_t ntable, fpm_stream_t stream);
int fpm_linear_scale_f64(void* field, const fpm_mesh* mesh, int64_t pitch2, const double* table,
                         int64_t ntable, fpm_stream_t stream);

/* delta_k *= table[index[mode]] with index an int32 array of one entry per
 * half-spectrum mode ([n0][n1][n2/2+1], C order; every entry in [0, ntable)
 * is the caller's guarantee) and table holding sqrt(P(k)/V): Solver.linear
 * (fastpm/core.py:116-118) for a box whose sides differ, where the caller
 * tabulates P at the distinct |k| of the mesh. */
int fpm_linear_scale_indexed_f32(void* field, const fpm_mesh* mesh, int64_t pitch2,
                                 const double* table, int64_t ntable, const int32_t* index,
                                 fpm_stream_t stream);
int fpm_linear_scale_indexed_f64(void* field, const fpm_mesh* mesh, int64_t pitch2,
                                 const double* table, int64_t ntable, const int32_t* index,
                                 fpm_stream_t stream);

/* 2LPT source in real space from the six second derivatives
 * phi = {xx, yy, zz, xy, yz, zx}:
 * out = 3/7 (xx yy + yy zz + zz xx - xy^2 - yz^2 - zx^2).
 * Replaces the numpy products of fastpm/force/lpt.py:45-62. */
int fpm_lpt2_source_f32(const void* const* phi, void* out, int64_t ncells, fpm_stream_t stream);
int fpm_lpt2_source_f64(const void* const* phi, void* out, int64_t ncells, fpm_stream_t stream);

/* y += a * x over n elements (a in double).  Kick and drift updates,
 * fastpm/core.py:172 (P += fac F) and fastpm/core.py:179 (S += fac P). */
int fpm_axpy_f32(void* y, const void* x, int64_t n, double a, fpm_stream_t stream);
int fpm_axpy_f64(void* y, const void* x, int64_t n, double a, fpm_stream_t stream);

/* Consecutive kicks and a drift in one pass over n elements:
 * P += kicks[i] * F for i < nkick (0, 1 or 2, coefficients in double), then
 * S += drift * P, each update rounded exactly as fpm_axpy rounds it, so the
 * result equals the separate fpm_axpy calls bit for bit.  Replaces the
 * Kick, [Kick,] Drift sequence of a leapfrog step boundary
 * (fastpm/core.py:168-180 applied to the actions of core.py:261-288). */
int fpm_kick_drift_f32(void* S, void* P, const void* F, int64_t n, const double* kicks,
                       int nkick, double drift, fpm_stream_t stream);
int fpm_kick_drift_f64(void* S, void* P, const void* F, int64_t n, const double* kicks,
                       int nkick, double drift, fpm_stream_t stream);

/* y = a y + b x over n elements: the damped kick of the glass generator,
 * P = P + (af - ai) (F - P) (fastpm/glass.py:89-93). */
int fpm_axpby_f32(void* y, const void* x, int64_t n, double a, double b, fpm_stream_t stream);
int fpm_axpby_f64(void* y, const void* x, int64_t n, double a, double b, fpm_stream_t stream);

/* S = (S + a P) mod box_d on [n][3] arrays, results in [0, box_d): the
 * glass generator's drift (fastpm/glass.py:95-98). */
int fpm_drift_wrap_f32(void* S, const void* P, int64_t n, double a, const double* box,
                       fpm_stream_t stream);
int fpm_drift_wrap_f64(void* S, const void* P, int64_t n, double a, const double* box,
                       fpm_stream_t stream);

/* The transpose of fpm_kick_drift for the reverse pass of the differentiable
 * model: over n elements gP += drift * gS (the drift's VJP), then
 * gF += kicks[i] * gP for i < nkick (the kicks' VJPs), each rounded as
 * fpm_axpy, so the result equals the separate fpm_axpy calls bit for bit.
 * fresh != 0 means gF holds no gradient yet (it was just consumed by a
 * force's VJP): it then starts from zero and is not read.  With nkick = 0
 * gF is not touched.  The VJPs of Kick (fastpm/core.py:172, P += fac F) and
 * Drift (core.py:179, S += fac P) in the reverse order of a leapfrog step
 * boundary (core.py:261-288). */
int fpm_drift_kick_vjp_f32(const void* gS, void* gP, void* gF, int64_t n, const double* kicks,
                           int nkick, double drift, int fresh, fpm_stream_t stream);
int fpm_drift_kick_vjp_f64(const void* gS, void* gP, void* gF, int64_t n, const double* kicks,
                           int nkick, double drift, int fresh, fpm_stream_t stream);

/* y += table[counter[0] * stride + slot] * x, with the FP64 coefficient table
 * ([rows][stride]) and the step counter in device memory, so a captured graph
 * can replay the kick/drift updates of every step (fastpm/core.py:168-180
 * with the coefficients precomputed per step); a counter outside [0, rows)
 * leaves y unchanged.  fpm_counter_add adds v to counter[0] (a kernel, safe
 * inside graph capture). */
int fpm_axpy_table_f32(void* y, const void* x, int64_t n, const double* table,
                       const int32_t* counter, int rows, int stride, int slot,
                       fpm_stream_t stream);
int fpm_axpy_table_f64(void* y, const void* x, int64_t n, const double* table,
                       const int32_t* counter, int rows, int stride, int slot,
                       fpm_stream_t stream);
int fpm_counter_add(int32_t* counter, int32_t v, fpm_stream_t stream);

/* The table-driven form of fpm_kick_drift for a captured step graph: with
 * row = counter[0], P += table[row-1][close] F (the previous step's closing
 * kick; skipped for row 0), P += table[row][open] F, S += table[row][drift] P,
 * each rounded as fpm_axpy, so a replay equals the separate updates bit for
 * bit.  A counter outside [0, rows) leaves S and P unchanged.  The leapfrog
 * step boundary of fastpm/core.py:168-180,261-288 with the coefficients
 * precomputed per step. */
int fpm_kick_drift_table_f32(void* S, void* P, const void* F, int64_t n, const double* table,
                             const int32_t* counter, int rows, int stride, int close, int open,
                             int drift, fpm_stream_t stream);
int fpm_kick_drift_table_f64(void* S, void* P, const void* F, int64_t n, const double* table,
                             const int32_t* counter, int rows, int stride, int close, int open,
                             int drift, fpm_stream_t stream);

/* S = c0 dx1 + c1 dx2; P = c2 dx1 + c3 dx2; F = c4 dx1 + c5 dx2 over n
 * elements (dx2 may be NULL for first order).  fastpm/core.py:128-141. */
int fpm_lpt_combine_f32(const void* dx1, const void* dx2, void* S, void* P, void* F, int64_t n,
                        const double* coef, fpm_stream_t stream);
int fpm_lpt_combine_f64(const void* dx1, const void* dx2, void* S, void* P, void* F, int64_t n,
                        const double* coef, fpm_stream_t stream);

/* out[p][d] = Q_p[d] + S_p[d] (StateVector.X, fastpm/core.py:41-43). */
int fpm_positions_f32(const fpm_particles* part, const fpm_mesh* mesh, void* out,
                      fpm_stream_t stream);
int fpm_positions_f64(const fpm_particles* part, const fpm_mesh* mesh, void* out,
                      fpm_stream_t stream);

/* Shell-binned spectra with double accumulation, deterministic (fixed
 * reduction order, no atomics).  Bins: [kmin + b dk, kmin + (b+1) dk),
 * b < nbins; the zero mode is excluded; half-spectrum modes with
 * 0 < k_z index < n2/2 count twice.  out[b][0..4] = sum w, sum w |k|,
 * sum w Re(f1 f2*), sum w |f1|^2, sum w |f2|^2 (f2 == NULL: f2 = f1).
 * work holds fpm_power_work_len() doubles.
 * Replaces the estimator of fastpm/utils.py:77-164 and the
 * nbodykit FFTPower call of fastpm/main.py:81-92. */
int64_t fpm_power_work_len(const fpm_mesh* mesh, int nbins);
int fpm_power_bin_f32(const void* f1, const void* f2, const fpm_mesh* mesh, int64_t pitch2,
                      int nbins, double kmin, double dk, double* work, int64_t work_len,
                      double* out, fpm_stream_t stream);
int fpm_power_bin_f64(const void* f1, const void* f2, const fpm_mesh* mesh, int64_t pitch2,
                      int nbins, double kmin, double dk, double* work, int64_t work_len,
                      double* out, fpm_stream_t stream);

/* fpm_power_bin without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * monitoring loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The block partials live in a fixed device array of the library
 * (10 MiB, 1310720 doubles, held on every device the library runs on until
 * the process ends); a call needs fpm_power_work_len(mesh, nbins) doubles of
 * it, so up to 128 bins on any mesh (more on meshes below 2048 x 8192 modes),
 * and returns FPM_ERR_SCRATCH beyond that.  Same results as fpm_power_bin,
 * bit for bit.  It allocates nothing; its first call on a device looks the
 * array's address up (no allocation).  The stream must belong to the current
 * device.  All calls on one device share the array: calls that overlap on a
 * device (other streams or threads) race. */
int fpm_power_bin_noalloc_f32(const void* f1, const void* f2, const fpm_mesh* mesh,
                              int64_t pitch2, int nbins, double kmin, double dk, double* out,
                              fpm_stream_t stream);
int fpm_power_bin_noalloc_f64(const void* f1, const void* f2, const fpm_mesh* mesh,
                              int64_t pitch2, int nbins, double kmin, double dk, double* out,
                              fpm_stream_t stream);

/* ---- Reverse-mode (adjoint) kernels.  They replace the pmesh gradient
 * operators the reference's hand-written adjoints call
 * (fastpm/operators.py:68-98,140-193,218-253: readout_gradient,
 * paint_gradient, c2r_gradient, r2c_gradient, conjugated kernels). */

/* Paint with per-particle mass scale * weights[p * wstride] (the adjoint of
 * readout with respect to the mesh; readout_gradient's mesh output). */
int fpm_paint_weighted_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                           const void* weights, int64_t wstride, double scale, void* out,
                           int accumulate, fpm_stream_t stream);
int fpm_paint_weighted_f64(const fpm_particles* part, const fpm_mesh* mesh, int order,
                           const void* weights, int64_t wstride, double scale, void* out,
                           int accumulate, fpm_stream_t stream);

/* Position derivative of the window: gx[p][d] (+)= scale * sum_c w_c[p] *
 * sum_cells meshes[c](cell) dW(x_p - cell)/dx_d (weights == NULL or
 * weights[c] == NULL means 1).  Readout's position gradient
 * (readout_gradient) and paint's (paint_gradient). */
int fpm_window_grad_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                        const void* const* meshes, const void* const* weights,
                        const int64_t* wstrides, double scale, void* gx, int accumulate,
                        fpm_stream_t stream);
int fpm_window_grad_f64(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                        const void* const* meshes, const void* const* weights,
                        const int64_t* wstrides, double scale, void* gx, int accumulate,
                        fpm_stream_t stream);

/* acc (+)= M(k) in, or conj(M(k)) in when conj != 0, for the multiplier M of
 * `op` (the adjoint of a k-space multiplier is its conjugate,
 * fastpm/operators.py:13-22 diff_kernel(conjugate=True)). */
int fpm_kspace_accumulate_f32(void* acc, const void* in, const fpm_kop* op, int conj,
                              int accumulate, const fpm_mesh* mesh, int64_t pitch2,
                              fpm_stream_t stream);
int fpm_kspace_accumulate_f64(void* acc, const void* in, const fpm_kop* op, int conj,
                              int accumulate, const fpm_mesh* mesh, int64_t pitch2,
                              fpm_stream_t stream);

/* Scales the half-spectrum modes with 0 < k_z index < n2/2 by `mid` and the
 * k_z = 0 / n2/2 planes by `edge` (Hermitian bookkeeping of c2r_gradient /
 * r2c_gradient; fastpm/operators.py:94-96). */
int fpm_hermitian_scale_f32(void* field, double mid, double edge, const fpm_mesh* mesh,
                            int64_t pitch2, fpm_stream_t stream);
int fpm_hermitian_scale_f64(void* field, double mid, double edge, const fpm_mesh* mesh,
                            int64_t pitch2, fpm_stream_t stream);

/* The force's four-point gradient as a real-space stencil,
 * D_d f = [8 (f(x+h) - f(x-h)) - (f(x+2h) - f(x-2h))] / (12 h), h = box_d / n_d,
 * equal to c2r(i a_d(k) r2c(f)) with the order-1 kernel of
 * fastpm/force/kernels.py:17-24 (periodic wrap).
 * fpm_grad4_mesh: outs[d] = scale * D_d phi (three meshes, not aliasing phi).
 * fpm_grad4_div: out (+)= scale * sum_d D_d g[d]; D_d^T = -D_d, so scale = -1
 * gives the adjoint of fpm_grad4_mesh.  Used by the force's adjoint
 * (gravity_gradient, fastpm/operators.py:218-253). */
int fpm_grad4_mesh_f32(const void* phi, void* const* outs, const fpm_mesh* mesh, double scale,
                       fpm_stream_t stream);
int fpm_grad4_mesh_f64(const void* phi, void* const* outs, const fpm_mesh* mesh, double scale,
                       fpm_stream_t stream);
int fpm_grad4_div_f32(const void* const* g, void* out, const fpm_mesh* mesh, double scale,
                      int accumulate, fpm_stream_t stream);
int fpm_grad4_div_f64(const void* const* g, void* out, const fpm_mesh* mesh, double scale,
                      int accumulate, fpm_stream_t stream);

/* gphi[i] = gsource * d source / d phi_i for the 2LPT source of
 * fpm_lpt2_source (fastpm/operators.py:140-193). */
int fpm_lpt2_source_vjp_f32(const void* gsource, const void* const* phi, void* const* gphi,
                            int64_t ncells, fpm_stream_t stream);
int fpm_lpt2_source_vjp_f64(const void* gsource, const void* const* phi, void* const* gphi,
                            int64_t ncells, fpm_stream_t stream);

/* ---- Batched boxes: one step boundary and force of many independent boxes
 * in a few launches.  No counterpart in the reference, which runs one box
 * per process; the per-box arithmetic is the single-box path's
 * (fpm_kick_drift_table, fpm_bin_particles_f32, fpm_paint_tiled_f32,
 * fpm_fft_r2c_c2r_f32 without the spectrum, fpm_readout_tiled_grad4_f32:
 * the Kick, Kick, Drift and Force of fastpm/core.py:168-210), and the
 * results equal those calls bit for bit.  Each launch covers every box that
 * uses the same compiled kernel (the window order, the mesh's variant
 * fields, the transform lengths), with one workgroup per tile, particle span
 * or row group of every box in one index space, each box's arguments read
 * from a table in device memory.
 *
 * One box: float32 storage, a force mesh that is a multiple of 16 and at
 * least 16 + order + 3 cells per side, lattice or explicit particles. */
typedef struct fpm_batch_box {
  fpm_particles part;      /* part.s is the displacement S, updated by the drift */
  fpm_mesh mesh;           /* force mesh, with its variant word */
  int32_t order;           /* 2 CIC, 3 TSC */
  int32_t pad_;
  double mass;             /* paint mass per particle (1 / nbar) */
  fpm_kop common;          /* potential multiplier, e.g. 3/2 Om0 / k^2 [/ W^p] */
  int64_t pitch2;
  void* rho;               /* real mesh: the painted density */
  void* phi;               /* real mesh: the potential */
  void* work;              /* complex mesh of pitch2: the transform's work space */
  const void* tw[3];       /* twiddle tables of the three axes */
  void* ws;                /* tile workspace, fpm_tile_workspace_size bytes */
  int64_t ws_bytes;
  void* rho_out;           /* RHO [n] (stride rho_stride), or NULL */
  int64_t rho_stride;
  void* force_out;         /* F [n][3], stride force_stride (>= 3) per particle */
  int64_t force_stride;
  double rho_scale, force_scale;
  /* the step boundary before the force, as fpm_kick_drift_table: S = part.s,
   * P, F = force_out (force_stride 3), table [rows][stride] */
  void* P;
  const double* table;
  int32_t rows, stride, close, open, drift, pad2_;
} fpm_batch_box;

typedef struct fpm_batch fpm_batch;  /* host-side plan (opaque) */

/* Device bytes the plan of these boxes needs for its argument tables. */
int64_t fpm_batch_device_bytes(int nbox, const fpm_batch_box* boxes);

/* Builds the plan: checks every box, writes the argument tables to `dev`
 * (fpm_batch_device_bytes bytes of device memory, owned by the caller and
 * kept alive and unchanged while the plan is used) and returns the host-side
 * plan, or NULL with *status set.  A set-up call: it copies to the device
 * synchronously.  The boxes' buffers must stay where they are. */
fpm_batch* fpm_batch_create(int nbox, const fpm_batch_box* boxes, void* dev, int64_t dev_bytes,
                            int* status);

/* Enqueues, for every box: with counter != NULL the step boundary
 * P += table[row-1][close] F (skipped for row 0), P += table[row][open] F,
 * S += table[row][drift] P (row = counter[0]; a row outside [0, rows)
 * leaves S and P unchanged; applied by the binning's pass over the
 * particles, as fpm_kick_drift_rebin_f32 does), then the force: binning,
 * paint of rho (mass
 * per particle), phi = c2r(common * r2c(rho)), RHO = rho_scale
 * readout(rho) and F = force_scale readout(D phi) with the four-point
 * gradient stencil.  The binning rebins each box into the capacity layout
 * its previous step left (fpm_rebin_particles; the plan's first step places
 * the records exactly, as fpm_batch_create clears the layout's header), so
 * the boxes' workspaces must not be used for anything else while the plan
 * is.  Asynchronous, no allocation, graph-capture safe. */
int fpm_batch_step(const fpm_batch* plan, const int32_t* counter, fpm_stream_t stream);

void fpm_batch_destroy(fpm_batch* plan);

/* Growth functions of a matter + Lambda (+ curvature) universe, host-side
 * double precision.  For each a[i] writes 9 values to out[9 i ...]:
 * D1, dD1/dlna, d2D1/dlna2, D2, dD2/dlna, d2D2/dlna2, E, dE/dlna, Omega_m(a),
 * with D1 and D2 each normalised to 1 at a_normalize.
 * Replaces fastpm/background.py:4-256 (Perturbation / MatterDominated). */
int fpm_growth(double omega_m, double omega_lambda, double omega_k, double a_normalize,
               const double* a, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* FPM_H */
