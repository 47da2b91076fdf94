/* wt.h -- C ABI of the MI355X (gfx950) D3Q19 wind-tunnel solver library, libwt.so.
 *
 * Plain C: every entry point is extern "C", takes plain pointers and sizes,
 * never throws, and returns an int status (0 = ok, < 0 = error; the message is
 * in wt_last_error(), which is thread-local). Device buffers are allocated by
 * the caller (e.g. torch.empty(..., device="cuda")) and passed as raw device
 * pointers; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * All launches are asynchronous on `stream` unless stated otherwise.
 *
 * Reference interfaces each entry point replaces (RICKxxROLLING/wind-tunnel):
 *   lbm.py = server/lbm.py, app.py = server/app.py, voxelize.js = web/js/voxelize.js
 */
#ifndef WT_H
#define WT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Population storage forms. FP32 stores f_i as in lbm.py:404. FP16 and BF16
 * store the deviation g_i = f_i - w_i from the lattice weight (collision math
 * stays FP32). FP64 stores f_i in double and does all arithmetic in double. */
#define WT_FP32 0
#define WT_FP16 1
#define WT_BF16 2
#define WT_FP64 3

/* wt_step / wt_batch_step flags */
#define WT_FLAG_FORCES 1 /* fuse the momentum-exchange force into the last step */
#define WT_FLAG_MASS 2   /* fuse the FP64 mass reduction into the last step */
/* Leave the macroscopic field unwritten in every step of the call (mac keeps
 * the field of the last step that stored it). The field is an output only:
 * face cells read the shell snapshot, forces come from the populations, so
 * populations, shells and sums are the same bits with or without the flag. For
 * steps whose field nothing reads: no particle advection, slice or surface
 * pressure from it (the reference's Solver.step reads it once, after the last
 * of its steps: lbm.py:443-453). In a batch, tunnels whose gate byte is set
 * (wt_batch_gate_offset) still store it. No counterpart in the reference. */
#define WT_FLAG_NO_MAC 4
/* wt_batch_info.tile_order bits */
#define WT_BATCH_BACKWARDS 2
#define WT_BATCH_PERSISTENT 4

/* One tunnel's device state. n = nx*ny*nz; sizes come from wt_buffer_sizes().
 * f[0], f[1]: ping-pong populations, row-blocked SoA: each x-row
 *             r = y + ny*z holds its 19 planes of nx values in turn, so
 *             population i of cell idx = x + nx*r is element
 *             19*nx*r + nx*p[i] + x (the reference's plane i*n + idx,
 *             lbm.py:68, 118, cut into rows), with plane positions
 *             p = {0,1,2,3,6,9,12,4,7,8,5,10,13,14,11,15,16,17,18}
 *             (planes read by the same neighbouring row adjacent).
 *             19*n elements.
 * cls:        per-cell class byte built by wt_set_solid from the solid mask.
 * mac:        macroscopic field per cell (lbm.py:406): FP32 and FP64 forms
 *             float4 (ux, uy, uz, rho), 16 B; FP16 and BF16 forms four
 *             binary16 (ux, uy, uz, rho - 1), 8 B, each rounded to nearest
 *             even from the float32 value (rho - 1 formed in float32). The
 *             particle, slice and Cp kernels widen the binary16 values to
 *             float32 and take rho = 1 + (rho - 1) in float32.
 * shell[0,1]: ping-pong snapshot of the macroscopic field on the six layers
 *             next to the domain faces; face cells read the previous step's
 *             values from it.
 * partials:   per-tile FP64 reduction slots (force x,y,z and mass).
 * sums:       4 doubles on the device: fx, fy, fz, mass.
 * cpt, tile_order: launch shape of this tunnel's step kernels, chosen per
 *             size class by the autotuner; no result depends on it.
 *             cpt = cells per thread (1, 2 or 4 consecutive 256-cell slots per
 *             workgroup; 0 means 1). tile_order bits: bit 0 XCD-contiguous
 *             tiles, bit 1 sweep direction alternating with the buffer parity,
 *             bit 2 64-bit global population addresses even where 32-bit
 *             buffer offsets would do (the form large tunnels take by
 *             themselves). Every call reads the shape from the grid it is
 *             given: the library keeps no launch state between calls.
 * Alignment:  f, cls, mac and shell 16-byte aligned (the FP64 shell 32-byte),
 *             partials and sums 8-byte, as every device allocation is; calls
 *             reject grids with other buffers. */
typedef struct wt_grid {
  int32_t nx, ny, nz;
  int32_t form;
  int32_t cpt;
  int32_t tile_order;
  void* f[2];
  uint8_t* cls;
  void* mac;
  void* shell[2];
  double* partials;
  double* sums;
} wt_grid;

/* Per-tunnel physical parameters, p['uIn'], p['tau'], p['smag'] (lbm.py:446). */
typedef struct wt_params {
  float uin;
  float tau;
  float smag;
} wt_params;

/* Respawn box of the particle rake: p['seedX'], ['seedCy'], ['seedCz'],
 * ['seedHy'], ['seedHz'] (lbm.py:452-453). */
typedef struct wt_seed_box {
  float x, cy, cz, hy, hz;
} wt_seed_box;

/* ---- library ------------------------------------------------------------- */
const char* wt_last_error(void);
int wt_version(void);
/* Device name and total memory (replaces Warp's device query, app.py:56-59). */
int wt_device_info(char* name, int name_len, int64_t* total_bytes);

/* ---- sizes and admission -------------------------------------------------- */
/* Bytes of each buffer of wt_grid, in the order f (one buffer), cls, mac,
 * shell (one buffer), partials, sums (sizes[6]). Replaces the per-cell
 * footprint constant BYTES_PER_CELL (lbm.py:30) with exact sizes. */
int wt_buffer_sizes(int nx, int ny, int nz, int form, int64_t sizes[6]);
/* Total device bytes of one tunnel (the admission footprint, app.py:72-76). */
int64_t wt_state_bytes(int nx, int ny, int nz, int form);
/* Number of reduction tiles of a tunnel (partials holds 4 doubles each). */
int64_t wt_tile_count(int nx, int ny, int nz);
/* Cache policy wt_step uses for a tunnel of this size and storage form, by
 * the bytes of its two population buffers: 1 (above 2 GiB) streams the
 * populations' c_x = 0 planes and the macroscopic field past the caches,
 * 2 (within 224 MiB) keeps everything cached, 0 otherwise (population stores
 * non-temporal). Batched steps use 1 for FP32, 0 for the other forms. No counterpart in the reference;
 * it names the kernel variant a measurement refers to. */
int wt_cache_policy(int nx, int ny, int nz, int form);

/* ---- state ---------------------------------------------------------------- */
/* Equilibrium at rho = 1, u = (uin, 0, 0) in both population buffers, mac and
 * both shells. Replaces Solver.reset / init_kernel (lbm.py:415-417, 64-69). */
int wt_init(const wt_grid* g, float uin, void* stream);
/* Build the class bytes from a device solid mask of n bytes (0 fluid,
 * 1 static solid, 2 moving belt, other non-zero = solid). Replaces
 * Solver.set_solid (lbm.py:419-421). */
int wt_set_solid(const wt_grid* g, const uint8_t* solid_dev, void* stream);

/* ---- the hot path --------------------------------------------------------- */
/* nsteps fused D3Q19 steps starting from f[cur] (reads f[cur], writes
 * f[cur^1], and so on). The caller flips its own cur nsteps times.
 * WT_FLAG_FORCES and WT_FLAG_MASS apply to the last step, and the reduced
 * values are in g->sums when the stream reaches the end of this call;
 * WT_FLAG_NO_MAC applies to every step. Replaces the step loop of Solver.step
 * (lbm.py:442-447) over step_kernel (lbm.py:72-166). */
int wt_step(const wt_grid* g, int cur, int nsteps, const wt_params* p, int flags, void* stream);
/* wt_step without allocation of any kind, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. The force/mass
 * reduction slots live in a static device array of the library (8.4 MB per
 * device, held until the process ends) instead of g->partials, which may be
 * NULL; g->sums receives the sums as with wt_step. Tunnels of at most 2^26
 * cells (larger ones are rejected with an error). Calls that overlap on one
 * device share that array and race: serialize them (one stream, or events). */
int wt_step_noalloc(const wt_grid* g, int cur, int nsteps, const wt_params* p, int flags, void* stream);
/* One lattice step as wt_step(g, cur, 1, p, flags, stream), fused in the same
 * launch with the particle advection of the previous step (slot storage, as
 * wt_advect_slots; ids may be NULL when slot = id) sampled from mac_prev, the
 * macroscopic field that step wrote. mac_prev must not be g->mac, which this
 * step overwrites: callers alternate two mac buffers. Removes one kernel
 * boundary per step (lbm.py:442-453 launches the step and advect_kernel
 * separately); every result equals the two separate launches bit for bit.
 * flags: WT_FLAG_FORCES and WT_FLAG_MASS (the next step's advection reads the
 * field this step stores, so WT_FLAG_NO_MAC is rejected). */
int wt_step_advect(const wt_grid* g, int cur, const wt_params* p, int flags, const void* mac_prev, float* pos,
                   float* spd, const uint32_t* ids, int64_t num_p, float dt, float uin, float max_age,
                   uint32_t seed, const wt_seed_box* box, void* stream);
/* Momentum-exchange force on f[cur] into g->sums[0..2] (FP64, deterministic).
 * Replaces forces.zero_() + force_kernel (lbm.py:479-481, 169-197). */
int wt_forces(const wt_grid* g, int cur, void* stream);
/* Copy g->sums to host (synchronizes the stream). */
int wt_read_sums(const wt_grid* g, double out[4], void* stream);

/* ---- batched tunnels ------------------------------------------------------ */
/* Many resident tunnels advance one step each in ONE launch: one workgroup
 * per work item, tile k of tunnel t, where tunnel t's tiles hold
 * 256 * grids[t].cpt cells (each tunnel keeps the launch shape of its size
 * class). Replaces running one Solver.step per tunnel (lbm.py:442-447) for
 * the serving loop (app.py:150-202 generalised to many sessions).
 * total_tiles: work items of one step. tile_order: bit 1 (WT_BATCH_BACKWARDS)
 * sweeps the items backwards on odd parities (set by wt_batch_prepare; a caller
 * may clear it before wt_batch_step); bit 2 (WT_BATCH_PERSISTENT, set by the
 * caller) runs the persistent kernel: one grid of the device's full occupancy
 * whose workgroups stride over the items, instead of one workgroup per item.
 * Results do not depend on either bit. persist_blocks: workgroups of a
 * persistent launch (0, the default: the device's full occupancy; a caller may
 * set fewer). max_cpt:
 * the largest cells-per-thread of the batch's tunnels (1, 2 or 4), which picks
 * the kernel variant; set by wt_batch_encode and not to be changed. */
typedef struct wt_batch_info {
  int64_t total_tiles;
  int32_t max_chunks;
  int32_t form;
  int32_t ntunnels;
  int32_t tile_order;
  int32_t max_cpt;
  int32_t persist_blocks;
} wt_batch_info;
/* Device workspace bytes for a batch of ntunnels (descriptor tables and the
 * per-tunnel gate bytes). The workspace passed to wt_batch_prepare and
 * wt_batch_step must be 16-byte aligned (any device allocation is). */
int64_t wt_batch_workspace_bytes(int ntunnels);
/* Byte offset in the workspace of the batch's ntunnels gate bytes: with
 * WT_FLAG_NO_MAC, wt_batch_step stores the macroscopic field of tunnel k only
 * if gate byte k is non-zero (the encoded table sets every gate to 1). A
 * caller writes them in stream order before the step whose field it will read
 * (the serving loop: the sessions whose slice is due). No counterpart in the
 * reference. */
int64_t wt_batch_gate_offset(int ntunnels);
/* Upload the descriptor table of a batch: grids[k] with params[k], current
 * buffer cur[k]. All tunnels share one storage form. If sums_dev is not NULL
 * (device array of 2*ntunnels*4 doubles) the reduced sums of tunnel k after a
 * step of parity p land at sums_dev[(p*ntunnels + k)*4 ...], so one copy
 * fetches the whole batch; otherwise in grids[k].sums. Synchronizes `stream`. */
int wt_batch_prepare(const wt_grid* grids, const wt_params* params, const int* cur, int ntunnels,
                     double* sums_dev, void* workspace, wt_batch_info* info, void* stream);
/* The descriptor table wt_batch_prepare uploads, written into host memory
 * (wt_batch_workspace_bytes(ntunnels) bytes at host_image) with info filled
 * in; no device call. A caller that copies host_image to the workspace in
 * stream order (e.g. from pinned memory) can change a batch's membership
 * without synchronizing: the serving loop does, so that admitting or retiring
 * a session never drains the ticks in flight. No counterpart in the reference. */
int wt_batch_encode(const wt_grid* grids, const wt_params* params, const int* cur, int ntunnels,
                    double* sums_dev, void* host_image, wt_batch_info* info);
/* One step of every tunnel of the prepared batch; with parity p tunnel k
 * reads buffer cur[k]^p. flags as for wt_step (WT_FLAG_NO_MAC subject to the
 * gate bytes); sums land in each grid's sums. Kernel launches only: a caller
 * may capture it in a graph. */
int wt_batch_step(void* workspace, const wt_batch_info* info, int parity, int flags, void* stream);

/* ---- particles, slices, surface pressure ---------------------------------- */
/* RK2 particle advection with respawn (advect_kernel, lbm.py:233-263).
 * pos: float4 (x, y, z, age) per particle; spd: float per particle. */
int wt_advect(const wt_grid* g, float* pos, float* spd, int64_t num_p, float dt, float uin,
              float max_age, uint32_t seed, const wt_seed_box* box, void* stream);
/* Same as wt_advect, with lane j of the launch taking particle order[j]
 * (a permutation of 0..num_p-1, e.g. from wt_particle_order) so that a wave
 * samples neighbouring cells. Each particle's arithmetic and random stream
 * (keyed by its id) are unchanged, so results equal wt_advect's bit for bit. */
int wt_advect_ordered(const wt_grid* g, float* pos, float* spd, const uint32_t* order, int64_t num_p,
                      float dt, float uin, float max_age, uint32_t seed, const wt_seed_box* box,
                      void* stream);
/* Slot storage: slot j of pos/spd holds particle ids[j] (a permutation of
 * 0..num_p-1), e.g. after wt_particle_gather put the particles in cell order,
 * so a wave reads and writes contiguous slots. The random stream is keyed by
 * the id, so every particle's result equals wt_advect's bit for bit. No
 * counterpart in the reference (advect_kernel, lbm.py:233-263, in another
 * storage order). */
int wt_advect_slots(const wt_grid* g, float* pos, float* spd, const uint32_t* ids, int64_t num_p, float dt,
                    float uin, float max_age, uint32_t seed, const wt_seed_box* box, void* stream);
/* pos_out[j] = pos[order[j]], spd_out[j] = spd[order[j]], ids_out[j] =
 * ids[order[j]] (order[j] if ids is NULL); not in place. Puts slot storage in
 * the order wt_particle_order returns. No counterpart in the reference. */
int wt_particle_gather(const float* pos, const float* spd, const uint32_t* ids, const uint32_t* order,
                       int64_t num_p, float* pos_out, float* spd_out, uint32_t* ids_out, void* stream);
/* Device workspace bytes for wt_particle_order. */
int64_t wt_particle_order_workspace_bytes(int64_t num_p);
/* order[] = particle ids sorted by the index of the cell they are in (radix
 * sort on the device). No counterpart in the reference (a locality aid). */
int wt_particle_order(const wt_grid* g, const float* pos, int64_t num_p, uint32_t* order, void* workspace,
                      int64_t workspace_bytes, void* stream);
/* Quantize particles (pack_particles, lbm.py:266-275). */
int wt_pack_particles(const wt_grid* g, const float* pos, const float* spd, uint16_t* q16,
                      uint8_t* q8, int64_t num_p, void* stream);
/* Same from slot storage: slot j goes to position ids[j] of q16/q8, so the
 * output is in particle-id order (pack_particles, lbm.py:266-275). */
int wt_pack_particles_slots(const wt_grid* g, const float* pos, const float* spd, const uint32_t* ids,
                            uint16_t* q16, uint8_t* q8, int64_t num_p, void* stream);
/* One axis-aligned slice as (value byte, fluid byte) pairs (slice_kernel,
 * lbm.py:278-318; Solver.frame lbm.py:464-474). index must be in range. */
int wt_slice(const wt_grid* g, uint8_t* out, int axis, int index, int field, float uin, float lref,
             float vmin, float vmax, void* stream);
/* Surface pressure coefficient per vertex as int8 (cp_kernel, lbm.py:353-368). */
int wt_cp(const wt_grid* g, const float* verts, const float* norms, int8_t* out, int64_t nv,
          float uin, float cp_range, void* stream);

/* ---- checks ----------------------------------------------------------------- */
/* Per-cell link classification as the step kernel sees it (u64 per cell):
 * bits 0-1 cell class (0 interior fluid, 1 solid, 2 face); bit 8+i bounce-back
 * of population i; bit 27+i moving-belt term; bit 45+i link i counts toward the
 * force (i = 1..18). Mirrors lbm.py:81-84, 114-124, 185-188. */
int wt_link_classes(const wt_grid* g, uint64_t* out_dev, void* stream);
/* Self-test of the BF16 storage rounding: the device conversion the step
 * kernels store with against round-to-nearest-even in bit arithmetic, over all
 * 2^32 float bit patterns. counts_dev[0] (uint64) = differing non-NaN inputs,
 * counts_dev[1] = differing NaN inputs; both 0 when the two agree. No counterpart
 * in the reference, whose storage is FP32 only. */
int wt_selftest_bf16(uint64_t* counts_dev, void* stream);

/* ---- measurement ---------------------------------------------------------- */
/* Streaming device-to-device copy of `bytes` (multiple of 16, 16-byte aligned
 * buffers), one 16-byte element per thread; `blocks` is ignored. A plain
 * copy of known size: the reference point for achievable HBM bandwidth and for
 * scaling the FETCH_SIZE / WRITE_SIZE counters. No counterpart in the reference. */
int wt_stream_copy(void* dst, const void* src, int64_t bytes, int blocks, void* stream);

/* ---- geometry (host code) ------------------------------------------------- */
/* Voxelize a triangle soup given in lattice coordinates (9 floats per
 * triangle) into out_host[n] exactly as voxelize.js:5-115 does: ray parity
 * along +x, surface shell, flood-fill closure, ground layer y=0 tagged 2.
 * counts = {solid cells, frontal area, planform area} (before the ground tag). */
int wt_voxelize(const float* tris, int64_t ntris, int nx, int ny, int nz, int ground,
                uint8_t* out_host, int64_t counts[3]);
/* Union of ellipsoids (k rows of cx, cy, cz, ax, ay, az; a cell centre is
 * inside if sum(((p - c)/a)^2) <= 1), clipped to [1, dim-2], closed with the
 * voxelize.js flood fill, ground layer y=0 tagged 2 when ground != 0.
 * Host-side geometry for analytic bodies. */
int wt_ellipsoids(const double* ell, int k, int nx, int ny, int nz, int ground, uint8_t* out_host,
                  int64_t counts[3]);

#ifdef __cplusplus
}
#endif

#endif /* WT_H */
