/*
 * phonofdtd — C ABI of the MI355X (gfx950) FDTD time-stepping library that
 * replaces the step loop of phonomena's solver plugins.
 *
 * Every entry point returns int: 0 on success, a negative PF_ERR_* code on
 * failure (never throws, never aborts).  pf_last_error() gives the message of
 * the last failure on the calling thread.  A context is not thread-safe: one
 * host thread drives it.  The library owns all device memory; callers own the
 * host buffers they pass (plain pointers and element counts).
 *
 * Field arrays use the reference shapes and C order (z fastest), as in
 * phonomena/simulation/grid.py:88-98:
 *   ux (nx-1, ny, nz)   uy (nx, ny-1, nz)   uz (nx, ny, nz-1)
 * where nx, ny, nz are node counts (cells + 1 per axis).
 */
#ifndef PHONOFDTD_H
#define PHONOFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 2

#define PF_OK 0
#define PF_ERR_ARG (-1)     /* invalid argument or shape                 */
#define PF_ERR_HIP (-2)     /* HIP runtime error (message has details)   */
#define PF_ERR_STATE (-3)   /* call not valid in the current state       */
#define PF_ERR_NOMEM (-4)   /* device or host allocation failed          */
#define PF_ERR_RANGE (-5)   /* step count beyond the uploaded source     */

/* Arithmetic of the sweep. */
#define PF_FP64 0   /* FP64, the reference's expression order (parity mode);
                       when every spacing is a power of two the tiled sweep
                       multiplies by the exact reciprocals instead of
                       dividing, which gives the same bits                   */
#define PF_FP32 1   /* FP32 fields, pre-folded coefficients (production)      */
#define PF_FP32C 2  /* FP32 fields carried as hi+lo pairs: compensated
                       (TwoSum/Kahan) time integration for long runs         */

/* Sweep kernel.  PF_KERNEL_AUTO: PF_KERNEL_TILED2 for PF_FP32 meshes of at
   least PF_TILED2_AUTO_NODES nodes (>= 8 per axis), else PF_KERNEL_TILED;
   PF_KERNEL_NAIVE beyond 16 materials or a 4 GiB time level. */
#define PF_KERNEL_AUTO 0
#define PF_TILED2_AUTO_NODES (1LL << 24)
#define PF_KERNEL_NAIVE 1   /* one thread per node, stresses recomputed     */
#define PF_KERNEL_TILED 2   /* LDS-staged (y,z) tiles, register march in x  */
#define PF_KERNEL_STRIP 3   /* barrier-free: each wavefront owns a strip of
                               y rows in registers, z neighbours by DPP     */
#define PF_KERNEL_TILED2 4  /* two-step tiled sweep (PF_FP32 / PF_FP32C):
                               steps taken in pairs, levels n+1 and n+2 in
                               one pass over memory; an odd step runs the
                               one-step tiled sweep                         */

/* Field selectors for pf_get_field / pf_set_field. */
#define PF_UX 0
#define PF_UY 1
#define PF_UZ 2
#define PF_UX_OLD 3   /* previous level, as grid.ux_old after time_step */
#define PF_UY_OLD 4
#define PF_UZ_OLD 5

typedef struct pf_ctx pf_ctx;
typedef struct pf_batch pf_batch;

/*
 * Everything one mesh needs.  Replaces the state that
 * BaseSolver.init (phonomena/simulation/base_solver.py:194-222) builds:
 * the Grid spacings (grid.py:112-127), the Material arrays
 * (material.py:48-63) and dt (material.py:80-93).
 */
typedef struct pf_desc {
    int32_t nx, ny, nz;        /* node counts, each >= 4                       */
    int32_t precision;         /* PF_FP64 | PF_FP32 | PF_FP32C                  */
    int32_t kernel;            /* PF_KERNEL_*                                   */
    int32_t n_materials;       /* 1..256                                        */
    const double *fdx, *fdy, *fdz;   /* full spacings, nx-1 / ny-1 / nz-1       */
    const double *sdx, *sdy, *sdz;   /* staggered spacings, nx-2 / ny-2 / nz-2  */
    /* n_materials rows of 13: C00 C01 C02 C10 C11 C12 C20 C21 C22 (the normal
       3x3 block of the 6x6 stiffness), C33 C44 C55 (shear diagonal), rho.   */
    const double *materials;
    const uint8_t *ids;        /* nx*ny*nz material index per node, or NULL:
                                  rasterized on the device from the inclusion
                                  list below                                   */
    double dt;
    double dt2;                /* dt**2 exactly as the host computed it         */
    const double *mur;         /* 12 Mur coefficients: faces x=max, y=0, y=max,
                                  z=max (in that order), each (ux, uy, uz)      */
    const double *sponge_x, *sponge_y, *sponge_z;  /* per-axis damping terms
                                  (nx / ny / nz) or all NULL for none          */
    const double *wave;        /* FP64 source table w(tt), tt = 0..wave_len-1  */
    int64_t wave_len;          /* also the step budget of the context          */
    int32_t probe[6];          /* surface-u_z probe grid: i0 si ni j0 sj nj
                                  (ni == 0: no probes)                          */
    int32_t record_every;      /* probe frame after every k-th step            */
    int32_t probe_capacity;    /* ring size in frames                          */
    int32_t tile_j;            /* rows per block (tiled: 4/8/16 with tile_z
                                  64, 8/16/32 with tile_z 32; two-step: 16
                                  with tile_z 64, 32 with tile_z 32) or per
                                  wavefront (strip: 4/8); 0 = default          */
    int32_t tile_x;            /* x planes per block (0 = automatic: sized so
                                  the blocks fill whole rounds of the CUs)    */
    int32_t tile_z;            /* tiled: lanes along z per tile row, 64 (one
                                  row per wavefront) or 32 (two rows per
                                  wavefront).  tile_j = tile_z = 0 with the
                                  one-step tiled sweep: pf_create times each
                                  built shape on the new (zero) fields and
                                  keeps the fastest (identical bits); else 0
                                  = the one with fewer idle lanes.  A two-
                                  step context with tile_j = tile_z = 0
                                  measures the shape of its one-step sweep
                                  (odd and single steps) the same way         */
    /* Material map from an inclusion list (used when ids is NULL), the
       device form of Grid.inclusionIndices + Material.setConstants
       (grid.py:158-175, material.py:55-63): node (i, j, k) takes index 1
       when some inclusion m has sqrt((x[i]-tx)^2 + (y[j]-ty)^2) < r (FP64,
       the reference's order) and k < K_m, else index 0. */
    const double *node_x;      /* nx node coordinates along x                  */
    const double *node_y;      /* ny node coordinates along y                  */
    int32_t n_targets;         /* inclusions                                   */
    const double *targets;     /* n_targets x 4: tx, ty, r (the reference's
                                  float32 target record, widened), K_m = z
                                  nodes covered (0..nz, from the reference's
                                  z test)                                      */
    /* Free surface (k = 0).  0: the reference's, bit for bit, including its
       precedence in the u_z update (base_solver.py:514-516), where T3[..,1]
       enters unscaled and only T3[..,0] (= 0) is divided by fdz.  1: no
       counterpart in the reference, for physical spacings: the difference
       is divided, (T3[..,1] - T3[..,0]) / fdz, as in the interior update
       (:461-463).  With fdz = 1 both give the same bits; on a uniform grid
       the fixed scheme's fields do not depend on the spacing (with dt
       scaled alike), the reference's diverge at physical spacings.  The
       surface spacings stay the reference's (its first-only and swapped
       spacings at k = 0 are exact no-ops on uniform grids). */
    int32_t surface;
} pf_desc;

/* Library / device identification. */
int pf_abi_version(void);
const char *pf_last_error(void);
int pf_device_info(char *name, int32_t cap, int32_t *cu_count, double *hbm_gib);

/* Context life cycle: upload mesh, material, source; zero fields. */
int pf_create(const pf_desc *desc, pf_ctx **out);
int pf_destroy(pf_ctx *ctx);

/*
 * Advance nsteps time steps (the loop body of BaseSolver.run,
 * base_solver.py:245-260: source, stresses, free surface, displacements,
 * Mur faces, rotation).  Asynchronous on the context's stream.
 */
int pf_step(pf_ctx *ctx, int64_t nsteps);
int pf_sync(pf_ctx *ctx);
int pf_steps_done(pf_ctx *ctx, int64_t *out);
/*
 * The sweep a context runs (PF_KERNEL_AUTO resolved) and its tile: rows
 * along y and lanes along z per tile (the two-step tile for
 * PF_KERNEL_TILED2; 0 for the naive sweep).  No counterpart in the
 * reference, which has one solver loop; for reports and benchmarks.
 */
int pf_kernel_info(pf_ctx *ctx, int32_t *kernel, int32_t *tile_j, int32_t *tile_z);

/* Fields in reference shapes, converted to/from double. */
int pf_get_field(pf_ctx *ctx, int32_t which, double *host, int64_t count);
/*
 * pf_get_field without any device allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it):
 * pf_get_field stages the packed FP64 field in a buffer it allocates on
 * first use, this variant in one static array per device of
 * PF_NOALLOC_FIELD_MAX doubles (32 MiB), held until the process ends.
 * Fields with more elements are refused (PF_ERR_ARG).  Every call on a
 * device shares that device's array: calls that overlap on one GPU (from
 * several host threads) race.  Same result as pf_get_field.
 */
#define PF_NOALLOC_FIELD_MAX (4LL * 1024 * 1024)
int pf_get_field_noalloc(pf_ctx *ctx, int32_t which, double *host, int64_t count);
/*
 * Asynchronous snapshots of the current level, for the HDF output path
 * (Writer, base_solver.py:72-169, put :256-260): ux, uy, uz in the
 * reference's shapes as FP64, back to back, in pinned host memory.
 * pf_snapshot_open allocates `slots` (1..16) staging frames on the device
 * and in pinned host memory.  pf_snapshot(slot) queues the packing of the
 * current level on the context's stream and its copy to the host on a
 * second stream, and returns at once; steps issued afterwards run while the
 * copy is in flight.  pf_snapshot_wait(slot) waits for the copy and gives
 * the host frame and the step count it was taken at; the frame stays valid
 * until pf_snapshot_release(slot) frees the slot for the next snapshot.
 */
int pf_snapshot_open(pf_ctx *ctx, int32_t slots);
int pf_snapshot(pf_ctx *ctx, int32_t slot);
int pf_snapshot_wait(pf_ctx *ctx, int32_t slot, double **host, int64_t *steps_done);
int pf_snapshot_release(pf_ctx *ctx, int32_t slot);

/* The per-node material index map (nx*ny*nz bytes, z fastest), as uploaded
   or rasterized; Material.ids of material.py:48-63. */
int pf_get_ids(pf_ctx *ctx, uint8_t *host, int64_t count);
int pf_set_field(pf_ctx *ctx, int32_t which, const double *host, int64_t count);

/* Probe frames recorded so far (oldest first, at most capacity). */
int pf_get_probes(pf_ctx *ctx, double *host, int64_t cap_frames, int64_t *nframes);

/* Energy audit of the newest level, summed in FP64 in a fixed order. */
int pf_energy(pf_ctx *ctx, double *out);

/*
 * Spectral analysis (replaces the numpy transforms of analysis.spectrum,
 * phonomena/simulation/analysis.py:44-96): a batch of 1-D DFTs along the
 * contiguous axis in FP64, first nf frequencies,
 *   out[b, k] = scale * sum_t (in[b, t] * window[t]) exp(-2 pi i k t / n),
 * in (batch, n) row-major, out (batch, nf).  in_im and window may be NULL
 * (real input, no window).  Host arrays; runs on the current device.
 * O(n log n) per trace: radix-2 FFT for a power of two n <= 8192, the
 * chirp-z (Bluestein) form otherwise; n <= 2^23, batch <= 65535.
 */
int pf_dft(const double *in_re, const double *in_im, int64_t batch, int64_t n, int64_t nf,
           const double *window, double scale, double *out_re, double *out_im);

/*
 * Timing: run nsteps steps on the context's stream with HIP events around
 * the sweep launch of every fourth step (with PF_KERNEL_TILED2, of every
 * fourth pair of steps, whose sweep is the two-step pass; a wider stride
 * where that would bracket more than eight), the first included, and one
 * after the last step; returns the mean duration per step of the sweeps
 * (sweep_ms) and of everything else (face_ms: faces, probes, launch gaps;
 * with PF_KERNEL_TILED2 also the band, the level n+2 cells next to the x
 * and y faces that the two-step pass leaves).  Each event delays the
 * launch after it by a few microseconds, hence the stride.  The steps and
 * their results are those of pf_step.  The first call on a context creates
 * its timing events (33), which later calls reuse and pf_destroy frees.
 */
int pf_profile(pf_ctx *ctx, int64_t nsteps, double *sweep_ms, double *face_ms);

/*
 * Batches (no counterpart in the reference, whose solver steps one mesh):
 * many contexts of one precision stepped together.  A batched step is three
 * launches whatever the member count: one persistent sweep over a combined
 * tile work-list of all members, then all members' boundary faces, then
 * their probes and step counters; the step loop is captured in a HIP graph.  Each member
 * stops at its own step budget (wave_len) while the others go on, and its
 * fields are bit-identical to stepping it alone.  While a context is a
 * member, pf_step/pf_set_field/pf_profile on it fail (PF_ERR_STATE) and
 * pf_get_field/pf_get_probes/pf_energy wait for the batch's work.
 */
int pf_batch_create(int32_t precision, pf_batch **out);
int pf_batch_add(pf_batch *b, pf_ctx *ctx);      /* same precision, <= 16
                                                    materials, all members
                                                    with or all without an
                                                    absorbing layer          */
int pf_batch_remove(pf_batch *b, pf_ctx *ctx);
int pf_batch_step(pf_batch *b, int64_t nsteps);  /* asynchronous            */
int pf_batch_sync(pf_batch *b);
/* Tile shape of the batched sweep: rows (tile_j) x lanes along z (tile_z:
   16x64, 8x64, 32x32, 16x32, 8x32; 0 = default), planes per x chunk
   (tile_x, 0 = 32), and whether steps are replayed from a HIP graph. */
int pf_batch_set_tiles(pf_batch *b, int32_t tile_j, int32_t tile_z, int32_t tile_x, int32_t use_graph);
/* Autotuner: times every candidate shape over trial_steps real steps (all
   shapes give the same bits, so the members simply advance) and keeps the
   fastest; returns it and its time per batched step.                     */
int pf_batch_autotune(pf_batch *b, int32_t trial_steps, int32_t *tile_j, int32_t *tile_z, int32_t *tile_x,
                      double *ms_per_step);
int pf_batch_destroy(pf_batch *b);

/*
 * Diagnostics (no counterpart in the reference).  Streams `reads` input and
 * `writes` output float arrays of n elements each through the device `reps`
 * times with a coalesced one-pass kernel (each block a contiguous run of
 * float4 elements, all of a thread's loads issued before their use) in a few
 * launch shapes (1 or 2 float4 per stream and thread, 256 or 512 threads,
 * plain or non-temporal accesses) and returns the best achieved GB/s and its
 * mean ms per pass: the practical HBM bandwidth of that mix on this device,
 * and a kernel of known traffic for calibrating hardware counters.  Built
 * mixes (reads/writes): 1/1, 2/1, 6/3, 7/3.  Allocates its own buffers.
 */
int pf_stream_probe(int32_t reads, int32_t writes, int64_t n, int32_t reps, double *gbs, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* PHONOFDTD_H */
