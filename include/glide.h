/*
 * glide.h -- C ABI of the MI355X backward-LPDM step engine (libglide_hip.so).
 *
 * Plain C: pointers, sizes and POD structs only. Every pointer named "device"
 * is HIP device memory (for example a torch tensor's data_ptr() on ROCm); the
 * library never frees caller memory. Streams are hipStream_t passed as void*
 * (NULL = the default stream). Every entry point returns GLIDE_OK (0) or a
 * negative status; glide_last_error() gives the message of the last failure on
 * the calling thread. Status mapping used by the Python layer:
 *   GLIDE_ERR_INVALID -> ValueError, GLIDE_ERR_MEMORY -> MemoryError,
 *   GLIDE_ERR_DEVICE  -> RuntimeError.
 *
 * Reference interfaces replaced (mrghg/glide, paths under src/lpdm/):
 *   glide_step                 HannaScheme.step static branch + the driver's
 *                              active mask / time bin / escape kill
 *                              (turbulence/hanna.py:580-633, 855-1068;
 *                               main.py:1396-1401, 1459-1477)
 *   glide_footprint_keys       FootprintGridder.accumulate index math
 *                              (footprint_gridder.py:156-196)
 *   glide_footprint_accumulate FootprintGridder.accumulate scatter_add_
 *                              (footprint_gridder.py:197)
 *   GLIDE_F_FIXED +            the same scatter_add_, fused into the step as
 *   glide_fixed_finalize       an exact fixed-point sum (footprint_gridder.py:197)
 *   glide_resample_hour        ArcoEra5ZarrReader._resample_hour_to_agl
 *                              (met_reader.py:641-693; vertical_grid.py:166-316)
 *   glide_build_window         HannaScheme window stacks + raw met staging
 *                              (hanna.py:747-853, 1694-1830)
 *   glide_column_cape          compute_lcl_lnb_cape per met column
 *                              (convection/emanuel.py:282-399)
 *   glide_convect_columns      EmanuelReducedConvection.maybe_convect trigger +
 *                              closure (convection/emanuel.py:629-679)
 *   glide_convect              EmanuelReducedConvection matrix build +
 *                              _redistribute_particles, backward
 *                              (convection/emanuel.py:429-506, 720-805)
 *   glide_graph_*              HannaScheme._maybe_graph_compile CUDA-graph
 *                              capture of the step core (hanna.py:1117-1160)
 */
#ifndef GLIDE_H
#define GLIDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLIDE_ABI_VERSION 8

#define GLIDE_OK 0
#define GLIDE_ERR_INVALID (-1)
#define GLIDE_ERR_MEMORY (-2)
#define GLIDE_ERR_DEVICE (-3)

/* Met storage precision of a packed window. */
#define GLIDE_MET_F32 0
#define GLIDE_MET_BF16 1

/* GlideStepParams.flags */
#define GLIDE_F_MEANDER 0x1        /* Maryon meander OU (hanna.py:1028-1054) */
#define GLIDE_F_TL_FLOORS 0x2      /* FLEXPART T_L floors 10/10/30 s (hanna.py:1560-1563) */
#define GLIDE_F_SL_OVERRIDE 0x4    /* surface-layer override (hanna.py:1542-1554) */
#define GLIDE_F_EMIT 0x8           /* write footprint keys/values + escape kill */
#define GLIDE_F_MASK_INPUT 0x10    /* active mask given explicitly (GlideParticles.active) */
#define GLIDE_F_NO_TURB 0x20       /* skip the turbulence body (advection only) */
#define GLIDE_F_RELEASE_TERMINUS 0x40 /* lo_off holds each particle's release terminus te (window end -
                                         backward length): active iff r_off >= cursor > te, so a release
                                         is stepped exactly as when it runs alone (no counterpart in the
                                         reference, whose batch terminus gives non-earliest releases one
                                         more step, main.py:1383) */
#define GLIDE_F_PRESCRIBED 0x100   /* verification mode: sigma_u = sigma_v = sigma_w and T_L come from the
                                      support record (sigma_w_ft, T_Lw_ft) at every height, the column
                                      turbulence's free-troposphere branch taken everywhere; the engine-level
                                      counterpart of driving the OU primitives with given (sigma_w, T_L)
                                      (gpu_engine.py:345-507) */
#define GLIDE_F_FIXED 0x80         /* with GLIDE_F_EMIT: add each particle's residence to a
                                      fixed-point accumulator in the step (GlideEmit.fixed) instead of
                                      writing keys/values; replaces the reference's scatter_add_
                                      (footprint_gridder.py:197) with an order-free exact sum */

/*
 * Met grid and the per-level AGL ladder (metadata.lon/lat/level,
 * met_reader.py:85-99). lon0/lon1/lat0/lat1 are the first and last grid
 * coordinates (latitude may descend). levels: device float[nz], ascending.
 */
typedef struct GlideGrid {
    int32_t nx, ny, nz, reserved;
    double lon0, lon1, lat0, lat1;
    const float* levels;
} GlideGrid;

/*
 * One packed met window on the device (built by glide_build_window). Cells
 * are column-major, cell = (y * nx + x) * nz + level: a column's levels are
 * contiguous, so particles binned by column gather from one short run of
 * memory whatever their levels.
 *   GLIDE_MET_BF16: one 16 B record of 8 bfloat16 per cell,
 *     wind: u_s v_s w_s u_e v_e w_e 0 0
 *     aux : sigma_w_ft T_Lw_ft rho drho_dz sigma_mu sigma_mv 0 0
 *   GLIDE_MET_F32: two float planes over the cell index, plane A (16 B per
 *     cell) then plane B (8 B per cell, from byte ncell * 16), 24 B per cell:
 *     wind: A = u_s v_s u_e v_e, B = w_s w_e
 *     aux : A = sigma_w_ft T_Lw_ft rho drho_dz, B = sigma_mu sigma_mv
 *     (a gather's 16 B and 8 B entries of one column's levels share fewer
 *     cache lines than 32 B records would)
 *   surf: [ny][nx] records of 12 float: blh sp ustar shf t0 (hour start), same
 *     (hour end), 0 0, in both precisions.
 * Buffer sizes: wind and aux ncell * 16 B (bfloat16) or ncell * 24 B (float),
 * surf ny * nx * 48 B.
 */
typedef struct GlideWindow {
    const void* wind;
    const void* aux;
    const void* surf;
    int32_t dtype, reserved;
} GlideWindow;

/*
 * Particle state, structure of arrays (device). lon/lat/z may be strided
 * (pos_stride = 4 for the reference's (N, 4) [lon, lat, alt, weight] tensor,
 * 1 for separate arrays); all other arrays are contiguous.
 *   uid:     Philox counter word 0 (stable particle id).
 *   r_off:   release-time offset (s after batch start), we_off: window end
 *            offset, rel_idx: release index (release_generator.py:202-237).
 *   alive:   liveness (escape kill), active: explicit mask when
 *            GLIDE_F_MASK_INPUT is set (else active is computed from r_off,
 *            length and alive, main.py:1396-1401).
 *   lo_off:  per-particle earliest active cursor offset, r_off - backward
 *            length (releases of different lengths in one batch); NULL means
 *            r_off - p->length_s for every particle. With
 *            GLIDE_F_RELEASE_TERMINUS it holds the release terminus instead.
 * um/vm may be NULL when meander is off.
 *   cell:    (may be NULL) the precision layout: cell[i] = cy << 16 | cx, a
 *            point of the launch's met grid (nx, ny <= 65535), and lon/lat hold the offsets in
 *            degrees from that point, rlon = fma(cx, dlon, lon0) and rlat =
 *            fma(cy, dlat, lat0) in float32 (dlon = (lon1 - lon0) / (nx - 1),
 *            dlat likewise, rounded to float). A step moves the offsets and
 *            leaves each particle at its nearest grid point, so the offsets
 *            stay within about half a cell and keep float32 precision relative
 *            to the cell instead of to the coordinate. Taken by glide_step,
 *            glide_step_table, glide_convect, glide_reorder_keys and
 *            glide_permute_particles (pos_stride 1); the others reject it.
 */
typedef struct GlideParticles {
    int64_t n;
    int64_t pos_stride;
    float* lon;
    float* lat;
    float* z;
    float* up;
    float* vp;
    float* wp;
    float* um;
    float* vm;
    const float* weight;
    const float* r_off;
    const float* we_off;
    const int32_t* rel_idx;
    const uint32_t* uid;
    uint8_t* alive;
    const uint8_t* active;
    const float* lo_off;
    int32_t* cell;
} GlideParticles;

/*
 * Footprint geometry (FootprintGridder.__init__, footprint_gridder.py:50-115)
 * and the met domain of the escape kill (main.py:690-711).
 * The 32-bit key paths (keys/values, the sort accumulation, the release-binned
 * keys) need n_cells = n_r*n_t*n_z*n_y*n_x < 2^32 - 1; the fixed-point
 * accumulation takes any count (64-bit cells, per-release slabs).
 */
typedef struct GlideFootprintGrid {
    double lon_min, lon_max, lat_min, lat_max;
    const float* z_edges; /* device float[n_z + 1], strictly ascending */
    int32_t n_r, n_t, n_z, n_y, n_x, reserved;
    double dom_lon_min, dom_lon_max, dom_lat_min, dom_lat_max, dom_alt_max;
    /* Release-binned keys (device int32[n_r], may be NULL): when set, the
     * emitting kernels write keys r*(n_z*n_y*n_x) + (z*n_y + y)*n_x + x
     * (dropped: n_r*n_z*n_y*n_x) and t_bin[r] = release r's time-ago bin at
     * this step, for glide_footprint_accumulate_binned. Valid because every
     * particle of a release shares its window end (we_off), so within a step
     * the time bin is a property of the release: the keys are shorter (fewer
     * sort passes) and sort in the same order as the flat cell index. */
    int32_t* t_bin;
    /* Per-release slabs of the fixed-point accumulator (GLIDE_F_FIXED only; both
     * device arrays may be NULL): release r's cells are slab_off[r] ..
     * slab_off[r] + slab_nt[r] * n_z * n_y * n_x - 1, laid out [t][z][y][x], and
     * its time bins t >= slab_nt[r] are dropped; slab_cells is the accumulator's
     * total. Slabs sized by each release's length let a batch exceed the 2^32
     * cells of the flat key (2,688 releases x 24-120 hourly bins x 160 x 240).
     * NULL slab_off: the flat layout, slab_off[r] = r * n_t * n_z * n_y * n_x. */
    const int64_t* slab_off;
    const int32_t* slab_nt;
    int64_t slab_cells;
} GlideFootprintGrid;

/*
 * Per-step scalars and scheme constants.
 *   dt: this step's delta_s; alpha: the RK2-midpoint time weight;
 *   cursor_off: pre-step cursor (s after batch start); length_s: backward
 *   window length; seed/step_index: Philox key and counter word 1.
 *   block: threads per workgroup of the step kernel launch, a multiple of 64
 *   in [64, 512]; 0 = 256 (a tuning parameter: results do not depend on it).
 *   noise_hi: Philox counter word 3 of every draw (0 for a batch; a served
 *   request's word when its particles are convected on their own).
 */
typedef struct GlideStepParams {
    double dt, alpha, cursor_off, length_s;
    double substep_c, z_ubl_m, meander_timescale_s;
    uint64_t seed;
    int32_t step_index, max_substeps;
    int32_t flags, block;
    uint32_t noise_hi;
    int32_t reserved;
} GlideStepParams;

/* Step outputs (device): per-particle footprint key (n_cells = dropped) and
 * value (weight * dt), and the step's escape count (added to *escaped).
 * active (may be NULL): the step's count of active particles (its
 * particle-steps) is added to *active. In table mode (glide_step_table) both
 * counters point at per-row arrays indexed by the cursor.
 *
 * GLIDE_F_FIXED (flat keys only: GlideFootprintGrid.t_bin NULL): keys/vals
 * are not written (may be NULL); instead each particle's value v (the float
 * weight * dt) becomes the integer rint(v * 2^fixed_exp[r]) of its release r
 * and is added to fixed[key] (int64[n_cells], two's complement in 64-bit
 * words) by integer atomics after a per-wavefront sum over runs of equal
 * keys. Integer sums do not depend on order, so the result is bit-identical
 * whatever the launch shape or particle order; with fixed_exp chosen so every
 * release's whole residence stays below 2^62 (glide_fixed_finalize), the
 * integers hold each float contribution exactly in practice. fixed_copies
 * (>= 1; 0 = 1) copies of the n_cells accumulator lie back to back in fixed,
 * and workgroup b adds to copy b % fixed_copies: particles gathered in a few
 * cells (a point release) then spread their atomics over that many addresses
 * per cell instead of queueing on one. The cells' values are the sums over
 * the copies.
 *
 * active_copies (>= 1; 0 = 1): the step's active count is spread the same
 * way, workgroup b adding to active[b % active_copies] (table mode: the row's
 * active_copies counters start at active + row * active_copies); the count is
 * the sum of those counters. One counter per step takes one atomic from every
 * workgroup, which queue on its address (≈10 ns each). */
typedef struct GlideEmit {
    uint32_t* keys;
    float* vals;
    unsigned long long* escaped;
    unsigned long long* active;
    long long* fixed;
    const int32_t* fixed_exp;
    int32_t fixed_copies;
    int32_t active_copies;
} GlideEmit;

int glide_abi_version(void);
const char* glide_last_error(void);

/* Fused per-particle step: RK2 advection, Hanna sigma/T_L, drifts, substepped
 * exact-OU with reflection, meander, mask gate; optional footprint key/value
 * emission and escape kill (flags & GLIDE_F_EMIT). */
int glide_step(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
               const GlideWindow* w, const GlideFootprintGrid* fg, const GlideEmit* e,
               void* stream);

/* glide_step with the per-step scalars in device memory, for a captured step
 * body (replaces the reference's static-buffer graph replay,
 * hanna.py:1117-1160 with main.py:1396-1477): the step reads
 * table[row[0]] (device GlideStepParams rows, one per schedule step), adds its
 * escape count to e->escaped[row[0]], and the last
 * workgroup to finish advances row[0] by one. A cursor outside [0, n_rows)
 * steps no particle (with GLIDE_F_EMIT and no GLIDE_F_FIXED it writes
 * dropped keys), counts nothing and is left as it is. row points to GLIDE_ROW_WORDS
 * device int32: the cursor, then finished-workgroup counters that must be 0
 * at launch (they are left 0; glide_row_set sets them). tmpl is a host row
 * whose flags every step takes (the rows' flags fields are not read); the
 * argument checks use it, so the rows must meet the same constraints as tmpl
 * (dt >= 0, substep_c > 0, max_substeps >= 1; a row with max_substeps below 1
 * takes one substep). A call over no particles (s->n == 0) still advances an
 * in-range cursor by one. */
#define GLIDE_ROW_WORDS 66
int glide_step_table(const GlideStepParams* tmpl, const GlideStepParams* table, int32_t n_rows, int32_t* row,
                     const GlideParticles* s, const GlideGrid* g, const GlideWindow* w,
                     const GlideFootprintGrid* fg, const GlideEmit* e, void* stream);

/* Device cursor for glide_step_table: row[0] = value, the counters 0 (a
 * kernel, so it may sit inside a captured graph). */
int glide_row_set(int32_t* row, int32_t value, void* stream);

/*
 * Trajectory diagnostics (replace main.py:1565, the per-step
 * particles[:, :3].mean, and main.py:892-937, the hourly grid-mean wind).
 * Fixed-order float64 reductions: the same inputs give the same bits.
 *
 * glide_particle_means: mean lon, lat, z over all s->n particles (pos_stride
 * 1) into out[3r .. 3r+2]; r = row[0] - 1 when row is given (the step that
 * glide_step_table has just finished; nothing is written unless
 * 0 <= r < n_rows), else r = 0.
 * glide_field_means: out[c] = mean of fields[c * n_per_channel ..] over
 * n_per_channel elements, c < n_channels <= 4.
 * Both need glide_means_workspace_bytes() of device workspace.
 */
size_t glide_means_workspace_bytes(void);
int glide_particle_means(const GlideParticles* s, const int32_t* row, int32_t n_rows, double* out,
                         void* workspace, size_t workspace_bytes, void* stream);
int glide_field_means(const float* fields, int64_t n_per_channel, int32_t n_channels, double* out,
                      void* workspace, size_t workspace_bytes, void* stream);

/* FootprintGridder.accumulate index math on arbitrary inputs: positions
 * lon/lat/z (strided by pos_stride), active mask, weights, per-particle time
 * bin t_idx and release index (int64, as the reference passes them). */
int glide_footprint_keys(int64_t n, int64_t pos_stride, const float* lon, const float* lat,
                         const float* z, const uint8_t* active, const float* weight,
                         const int64_t* t_idx, const int64_t* rel_idx, double dt,
                         const GlideFootprintGrid* fg, uint32_t* keys, float* vals,
                         void* stream);

/* Workspace bytes glide_footprint_accumulate needs for n keys. */
size_t glide_footprint_workspace_bytes(int64_t n);

/* Deterministic residence accumulation: stable radix sort of (key, value)
 * by key, segmented reduction in a fixed order, one add per touched cell into
 * slab[n_cells] (float). comp (may be NULL) receives Neumaier compensation
 * terms, so slab + comp is the compensated sum. Keys equal to n_cells are
 * dropped. release_total (may be NULL): double[n_r] += per-release residence,
 * summed in a fixed order; cells_per_release = n_t*n_z*n_y*n_x. */
int glide_footprint_accumulate(const uint32_t* keys, const float* vals, int64_t n,
                               uint32_t n_cells, uint32_t cells_per_release, float* slab,
                               float* comp, double* release_total, void* workspace,
                               size_t workspace_bytes, void* stream);

/* glide_footprint_accumulate without a caller workspace and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The workspace is a static device array of
 * the library (19 MiB per GPU, held until the process ends) that serves up to
 * GLIDE_NOALLOC_MAX_PAIRS pairs; larger n is rejected (GLIDE_ERR_INVALID).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (concurrent streams or threads) race. Same results as
 * glide_footprint_accumulate, bit for bit. */
#define GLIDE_NOALLOC_MAX_PAIRS (1 << 20)
int glide_footprint_accumulate_noalloc(const uint32_t* keys, const float* vals, int64_t n, uint32_t n_cells,
                                       uint32_t cells_per_release, float* slab, float* comp,
                                       double* release_total, void* stream);

/* glide_footprint_accumulate for release-binned keys (GlideFootprintGrid.
 * t_bin): key k < n_r * cells_per_bin belongs to release r = k / cells_per_bin
 * and adds to slab cell (r * n_t + t_bin[r]) * cells_per_bin + k % cells_per_bin
 * (cells_per_bin = n_z*n_y*n_x); keys >= n_r * cells_per_bin are dropped. Same
 * order of additions as the flat keys, so the same result bit for bit.
 * release_total as in glide_footprint_accumulate. */
int glide_footprint_accumulate_binned(const uint32_t* keys, const float* vals, int64_t n, uint32_t n_r,
                                      uint32_t n_t, uint32_t cells_per_bin, const int32_t* t_bin, float* slab,
                                      float* comp, double* release_total, void* workspace,
                                      size_t workspace_bytes, void* stream);

/* Fixed-point footprint (GLIDE_F_FIXED) to float: slab[c] = (the sum over
 * the copies of fixed[k * n_cells + c]) * 2^-e rounded once to float32 (the correctly rounded exact sum of the float
 * contributions when each was held exactly), e = fixed_exp[c /
 * cells_per_release]; release_total[r] (may be NULL) = the exact integer sum
 * of release r's cells * 2^-e in float64 (overwritten, not added). The
 * exponent of a release whose residence is bounded by T (sum of its weights
 * times its backward length plus two steps) is 61 - ceil(log2 T). Workspace:
 * glide_fixed_workspace_bytes(n_r). */
size_t glide_fixed_workspace_bytes(int32_t n_r);
int glide_fixed_finalize(const long long* fixed, int32_t copies, int64_t n_cells, uint32_t cells_per_release,
                         const int32_t* fixed_exp, float* slab, double* release_total, void* workspace,
                         size_t workspace_bytes, void* stream);

/* glide_fixed_finalize for per-release slabs (GlideFootprintGrid.slab_off):
 * releases r0 .. r1-1, release r's slab_nt[r] * cells_per_bin cells from
 * slab_off[r] of fixed (copies back to back, slab_cells apart), to float
 * out[slab_off[r] - slab_off[r0] + c] (the range's slabs back to back when they
 * are contiguous in the accumulator); release_total[r] for r in [r0, r1) (may
 * be NULL) overwritten with the exact sums. max_nt bounds slab_nt (the launch
 * shape). Workspace: glide_fixed_workspace_bytes(r1). Converting a range of
 * finished releases lets a batch stream its footprints out per release. */
int glide_fixed_finalize_slabs(const long long* fixed, int32_t copies, int64_t slab_cells, const int64_t* slab_off,
                               const int32_t* slab_nt, uint32_t cells_per_bin, int32_t max_nt,
                               const int32_t* fixed_exp, int32_t r0, int32_t r1, float* out, double* release_total,
                               void* workspace, size_t workspace_bytes, void* stream);

/* Workspace bytes glide_sort_pairs needs for n pairs. */
size_t glide_sort_workspace_bytes(int64_t n);

/* Stable LSD radix sort of n 4-byte (key, value) pairs by the low key_bits of
 * the key (the engine's building block for the footprint and the particle
 * reordering; torch.sort(stable=True) of the reference's index math).
 * Outputs go to out_keys / out_vals (device). */
int glide_sort_pairs(const uint32_t* keys, const uint32_t* vals, int64_t n, int32_t key_bits,
                     uint32_t* out_keys, uint32_t* out_vals, void* workspace,
                     size_t workspace_bytes, void* stream);

/*
 * Particle reordering for gather locality and compaction of finished
 * particles (no counterpart in the reference, which keeps particles in release
 * order): keys[i] = met cell of particle i, column-major ((row * nx + column)
 * * nz + level); with n_releases > 1, plus rel_idx[i] * ncell (release-major:
 * a release's particles together, by cell; rel_idx clamped to the range), G =
 * n_releases groups, else G = 1; plus G*ncell (ncell = nz*ny*nx) for a particle
 * released further back than horizon_off (r_off < horizon_off: not active
 * before the cursor passes it); 2*G*ncell for a particle that can never be
 * active again (escaped, or its earliest active offset lo_off (r_off -
 * length_s) is after cursor_off); idx[i] = i. When live is not NULL, the number of particles that can be
 * active before the horizon is added to *live (zero it first). Sort with
 * glide_sort_pairs, then permute: those particles come first, then the ones
 * released further back. horizon_off below every r_off keeps one bucket.
 * When 2*G*ncell does not fit 32 bits (thousands of releases), the cell is the
 * column alone, ncell = ny*nx (a release's particles then sort by column).
 */
int glide_reorder_keys(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
                       double horizon_off, int32_t n_releases, uint32_t* keys, uint32_t* idx, uint32_t* live,
                       void* stream);

/* dst.array[j] = src.array[idx[j]] for every non-NULL array of the particle
 * state (structure of arrays, pos_stride 1, src and dst distinct). */
int glide_permute_particles(const GlideParticles* src, const GlideParticles* dst,
                            const uint32_t* idx, void* stream);

/*
 * Hourly terrain-following resample of one met hour (device arrays):
 *   src3 [c3][zp][ny][nx] float, height_agl [zp][ny][nx] float (negative below
 *   ground), agl_levels [za] double ascending, m_per_deg_lon [ny] double
 *   (clip(111320 cos(lat), 1, inf)), p_level_hpa [zp] double.
 *   dlat_deg/dlon_deg: uniform grid spacing (signed) for the terrain gradient,
 *   terrain [ny][nx] float.
 * Outputs: dst3 [c3][za][ny][nx] float; w (channel iw) slope-corrected with
 * u (iu), v (iv); p_out_hpa [za] double = horizontal mean of the regridded
 * level pressure.
 */
int glide_resample_hour(const float* src3, int32_t c3, int32_t zp, int32_t ny, int32_t nx,
                        const float* height_agl, const double* agl_levels, int32_t za,
                        const float* terrain, const double* m_per_deg_lon, double dlat_deg,
                        double dlon_deg, const double* p_level_hpa, int32_t iu, int32_t iv,
                        int32_t iw, float* dst3, double* p_out_hpa, void* stream);

/*
 * Per-hour unit conversions of raw source fields, in place, before
 * glide_resample_hour (replaces the reader's host-side conversions
 * ArcoEra5ZarrReader._compute_level_agl_m, met_reader.py:1052-1078,
 * _convert_vertical_velocity_to_m_s, 1268-1300, and _convert_shf_to_w_per_m2,
 * 1353-1385), with the reader's arithmetic:
 *   geopotential [zp][ny][nx], geopotential_sfc [ny][nx] float (m^2/s^2; may be
 *     NULL): height_agl [zp][ny][nx] = (geopotential - sfc) / g in float32, not
 *     clamped (negative below ground);
 *   w_ch >= 0: src3[w_ch] holds omega (Pa/s), replaced by w = -(R_d T / (g p))
 *     omega in float64 rounded to float32, T = src3[t_ch], p = p_level_pa[k];
 *   shf [ny][nx] float (may be NULL with GLIDE_SHF_UPWARD_W_M2): to upward-
 *     positive W/m^2 from downward-positive W/m^2 (negated) or from J/m^2
 *     accumulated over accumulation_s (divided, then negated), in float32.
 */
#define GLIDE_SHF_UPWARD_W_M2 0
#define GLIDE_SHF_DOWNWARD_W_M2 1
#define GLIDE_SHF_ACCUMULATED_J_M2 2
int glide_source_units(float* src3, int32_t c3, int32_t zp, int32_t ny, int32_t nx, int32_t w_ch, int32_t t_ch,
                       const double* p_level_pa, const float* geopotential, const float* geopotential_sfc,
                       float* height_agl, float* shf, int32_t shf_units, double accumulation_s, void* stream);

/*
 * Window build from two resampled hours (device):
 *   h3_s, h3_e: [5][nz][ny][nx] float (u, v, w, t, q) at hour start / end;
 *   h2_s, h2_e: [4][ny][nx] float (blh, sp, ustar, shf);
 *   p_level_hpa [nz] double (hour start), levels [nz] float.
 * Writes the packed records of GlideWindow (dtype selects float or bf16):
 * the window-midpoint free-troposphere sigma/T_L, rho and drho/dz, and the
 * meander sigma (coefficient * local std over the (2r+1)^2 stencil).
 */
int glide_build_window(const float* h3_s, const float* h3_e, const float* h2_s,
                       const float* h2_e, const double* p_level_hpa, const float* levels,
                       int32_t nz, int32_t ny, int32_t nx, double meander_coeff,
                       int32_t meander_radius, int32_t dtype, void* wind, void* aux, void* surf,
                       void* stream);

/*
 * Surface-parcel lift of every met column (compute_lcl_lnb_cape,
 * emanuel.py:282-399, which the reference applies to the bbox-mean column,
 * emanuel.py:593-627), float64 throughout. Per column: t = t_s (1 - alpha) +
 * t_e alpha and q likewise, clamped >= 0, from [nz][ny][nx] float fields with
 * level 0 the surface; pressure p_level_hpa [nz] double. Writes (each output
 * optional): i_lcl (first level with p < p_LCL, -1 when the parcel never
 * saturates), i_lnb (first level above the LCL with buoyancy <= 0, else
 * nz - 1; -1 without LCL), cape (J/kg), dtv (T_v excess at LCL+1, K; -1 when
 * off the column) and conv [4][ny][nx] float, the descriptors that
 * glide_convect_columns takes: (cape, levels[i_lcl], levels[i_lnb], dtv),
 * heights +inf when the index is -1.
 */
int glide_column_cape(const float* t_s, const float* t_e, const float* q_s, const float* q_e, double alpha,
                      const double* p_level_hpa, const float* levels, int32_t nz, int32_t ny, int32_t nx,
                      int32_t* i_lcl, int32_t* i_lnb, double* cape, double* dtv, float* conv, void* stream);

/*
 * Per-column convection descriptors for one met bracket (device):
 *   conv [4][ny][nx] float: CAPE (J/kg), LCL height, cloud-top height (m AGL),
 *   T_v excess at LCL+1 (K); t_s, t_e [nz][ny][nx] float resampled temperature
 *   at the bracket ends, lerped at alpha; p_level_hpa [nz] double; levels [nz].
 * A column fires when its LCL and cloud top lie on the ladder, top > LCL,
 * CAPE >= min_cape, T_v excess >= trigger and depth >= min_depth
 * (emanuel.py:637-650); then m_b = closure_c * rho_LCL * min(sqrt(2 CAPE), 5)
 * * 3600 s (emanuel.py:652-675). Writes i_lcl/i_lnb (-1 when not firing), m_b
 * and layer_mass [nz] = |dp|/g half brackets (emanuel.py:263-279).
 */
int glide_convect_columns(const float* conv, const float* t_s, const float* t_e, double alpha,
                          const double* p_level_hpa, const float* levels, int32_t nz, int32_t ny,
                          int32_t nx, double closure_c, double trigger_dtv_k, double min_cape_j_kg,
                          double min_cloud_depth_m, int32_t* i_lcl, int32_t* i_lnb, float* m_b,
                          float* layer_mass, void* stream);

/*
 * Per-column Emanuel convection (backward transpose sampling).
 * Column descriptors [ny][nx] (device): i_lcl, i_lnb (int32, -1 = no cloud),
 * m_b (float, kg/m^2 per event, already times the 3600 s interval).
 * layer_mass [nz] float (|dp|/g half-brackets), height [nz] float (ascending).
 * Active particles sample a destination level from the transposed, renormalised
 * column of the non-divergent mass-flux matrix of their nearest column; movers
 * get z uniform in [h[dest-1], h[dest]]. colsum_err (may be NULL): per-column
 * float64 max |row sum - column sum| / m (mass-conservation check). cloud (may
 * be NULL): device float[4*ny*nx] scratch (16 B aligned); when given, each
 * column's cloud totals (boundary-layer and detrainment-shape totals, CFL
 * scale, detrainment sum) are computed once into it before the particles
 * sample, instead of in every particle: the same arithmetic, the same result.
 * With GLIDE_F_EMIT in p->flags the footprint keys and the escape kill are
 * produced after the move, as glide_step does.
 */
int glide_convect(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
                  const int32_t* i_lcl, const int32_t* i_lnb, const float* m_b,
                  const float* layer_mass, double* colsum_err, float* cloud, const GlideFootprintGrid* fg,
                  const GlideEmit* e, void* stream);

/*
 * Launch timing, for measurement (no counterpart in the reference):
 * glide_step_timing_arm(n) makes the next n step-kernel launches of glide_step
 * and glide_step_table on the calling thread (outside graph capture) record
 * the start and end of the kernel itself (the launch's own events);
 * glide_step_timing_read(ms, n) waits for them, writes each one's duration in
 * milliseconds to ms[0 .. n) and returns how many launches were recorded
 * (arming again starts a new record).
 */
int glide_step_timing_arm(int32_t n);
int glide_step_timing_read(float* ms, int32_t n);

/*
 * Serving: many releases (requests), each with its own cursor, stepped
 * together. No counterpart in the reference, which steps one batch on one
 * cursor (main.py:1383-1477); each slot follows the reference's schedule for
 * a batch holding only its release (cursor from the window end down while
 * above the release terminus, main.py:1338-1389, bracket and alpha as
 * main.py:892-938), so a request's result does not depend on the others.
 *
 * GlideSlot (device array, one per slot): the host fills the first block on
 * admission (state = GLIDE_SLOT_RUNNING, t = window end, k = 0);
 * glide_slots_advance fills the per-step block for every running slot, or sets
 * state = GLIDE_SLOT_FINISHED once t <= term. Particle offsets (r_off,
 * we_off, lo_off) are relative to the slot's release time.
 */
#define GLIDE_SLOT_FREE 0
#define GLIDE_SLOT_RUNNING 1
#define GLIDE_SLOT_FINISHED 2

typedef struct GlideSlot {
    /* set on admission */
    double t;            /* cursor (s after the met epoch); starts at the window end */
    double term;         /* release terminus: window end - backward length */
    double rel_time;     /* release time: offsets are relative to it */
    int32_t state;
    int32_t k;           /* steps taken */
    uint32_t noise_hi;   /* Philox counter word 3 of the slot's particles (0 elsewhere): with
                            the particle ids it keeps concurrent requests' noise apart */
    int32_t convect;     /* per step: 1 when this is the first step of a met hour and
                            convection is on (its emission is left to glide_convect) */
    /* per step (glide_slots_advance) */
    double last_d;       /* length of the step taken last */
    float cursor_off, alpha, dt;
    int32_t noise_step;  /* Philox counter word 1: floor(t / dt) */
    int32_t bracket;     /* met hour floor(t / 3600) */
    int32_t stepping;    /* 1 when the slot takes a step this launch */
    /* counters (added by glide_step_slots) */
    unsigned long long escaped;
    unsigned long long active_steps;
} GlideSlot;

/* Met windows of consecutive brackets first_hour .. first_hour + n_hours - 1,
 * each as glide_build_window writes it, stride bytes apart (device). */
typedef struct GlideMetRing {
    const void* wind;
    const void* aux;
    const void* surf;
    int64_t wind_stride, aux_stride, surf_stride;
    int32_t first_hour, n_hours;
    int32_t dtype, reserved;
} GlideMetRing;

/* Advance every running slot one step of its schedule (dt = step length).
 * With GLIDE_SLOTS_CONVECT in flags a slot's step that is the first of a met
 * hour for it (its first step, or its bracket changed) is marked convect = 1:
 * glide_step_slots then steps its particles without emitting, and the caller
 * convects them and emits (glide_convect with the slot's scalars), the
 * reference's order (main.py:1446-1477). */
#define GLIDE_SLOTS_CONVECT 0x1
int glide_slots_advance(GlideSlot* slots, int32_t n_slots, double dt, int32_t flags, void* stream);

/* glide_step over a particle pool shared by the slots: wave_slot[j] is the slot
 * of particles 64j .. 64j+63 (-1: unused), each slot's particles lie in whole
 * such blocks. Each wavefront takes dt, alpha, cursor and noise step from its
 * slot and its met window from the ring (the slot's bracket); flags, seed and
 * the Hanna options from tmpl. Footprint keys use the slot as the release
 * index: slots holds fg->n_r entries, and a wave_slot entry outside
 * [0, fg->n_r) marks an unused block like -1; escape and active counts go to
 * the slot. Waves of non-stepping slots only write dropped keys. With
 * GLIDE_F_FIXED the residence goes to e->fixed as in glide_step, with
 * e->fixed_exp indexed by slot (and fg->slab_off, when given, placing each
 * slot's slab). Waves of a slot marked convect step without emitting or the
 * escape kill (GLIDE_F_FIXED only). */
int glide_step_slots(const GlideStepParams* tmpl, GlideSlot* slots, const int32_t* wave_slot,
                     const GlideParticles* s, const GlideGrid* g, const GlideMetRing* ring,
                     const GlideFootprintGrid* fg, const GlideEmit* e, void* stream);

/*
 * The Hanna (1982) / FLEXPART free functions element-wise over device float
 * arrays, evaluated with the step kernel's own float code and the window
 * build's closure (replaces the reference's free functions,
 * turbulence/hanna.py:142-189, 192-337, 340-426, for their unit tests and for
 * diagnostics). in/out are host arrays of device pointers (n floats each):
 *   CORIOLIS            in: lat (deg)                    out: f
 *   AIR_DENSITY         in: sp (Pa), t (K)               out: rho
 *   OBUKHOV             in: ustar, shf, t, sp            out: L (+inf when |kappa g H| <= 1e-10)
 *   CONVECTIVE_VELOCITY in: blh, shf, t, sp              out: w*
 *   IN_BL               in: z, blh, ustar, w*, h/L, lat  out: sigma_u, sigma_v, sigma_w, T_Lu, T_Lv, T_Lw
 *   SURFACE_LAYER_SIGMA_W in: z, ustar, L                out: sigma_w
 *   POTENTIAL_TEMPERATURE in: t, p (Pa)                  out: theta
 *   BRUNT_VAISALA       in: theta, dtheta/dz             out: N^2
 *   RICHARDSON          in: N^2, S^2                     out: Ri
 *   FT_DIFFUSIVITY      in: z, |S|, Ri                   out: K_z
 *   FT_SIGMA_TL         in: K_z, N^2                     out: sigma_w, T_Lw
 */
#define GLIDE_HANNA_CORIOLIS 1
#define GLIDE_HANNA_AIR_DENSITY 2
#define GLIDE_HANNA_OBUKHOV 3
#define GLIDE_HANNA_CONVECTIVE_VELOCITY 4
#define GLIDE_HANNA_IN_BL 5
#define GLIDE_HANNA_SURFACE_LAYER_SIGMA_W 6
#define GLIDE_HANNA_POTENTIAL_TEMPERATURE 7
#define GLIDE_HANNA_BRUNT_VAISALA 8
#define GLIDE_HANNA_RICHARDSON 9
#define GLIDE_HANNA_FT_DIFFUSIVITY 10
#define GLIDE_HANNA_FT_SIGMA_TL 11
int glide_hanna_eval(int32_t op, int64_t n, const float* const* in, float* const* out, void* stream);

/* HannaScheme._windowed_std (hanna.py:1786-1830): out[k][j][i] = population
 * std of field[k] over the edge-clipped (2r+1)^2 stencil around (j, i), the
 * meander stencil of glide_build_window. field, out: device float[nz][ny][nx]. */
int glide_windowed_std(const float* field, int32_t nz, int32_t ny, int32_t nx, int32_t radius, float* out,
                       void* stream);

/*
 * Step-body graph capture on a caller stream: glide_graph_begin starts stream
 * capture, glide_graph_end instantiates the captured work into a handle,
 * glide_graph_launch replays it, glide_graph_destroy frees it. Captured
 * kernels keep their arguments, so per-step scalars must live in device
 * memory that the caller refreshes between replays.
 */
int glide_graph_begin(void* stream);
int glide_graph_end(void* stream, void** handle);
int glide_graph_launch(void* handle, void* stream);
int glide_graph_destroy(void* handle);

#ifdef __cplusplus
}
#endif

#endif /* GLIDE_H */
