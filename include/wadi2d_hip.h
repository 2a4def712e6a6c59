/* C-ABI of the MI355X uniform-grid step (libwadi2d_hip.so).
 *
 * One handle holds one model's device arrays, as one backend object of the reference does
 * (wadi2d/solver/gpu/backend.py:53-203, the protocol shared with solver/cpu/backend.py:14-141). Every entry point
 * returns an int status: 0 ok, 1 invalid argument, 2 out of memory, 3 device error; w2d_last_error() gives the
 * message of the last failure on the calling thread. Host buffers are plain pointers with element counts.
 *
 * Arrays cross this boundary in the reference's state layout (wadi2d/solver/layout.py:1-52): tile-major,
 * 32 x 32 own cells inside a 2-cell halo, STRIDE 36, STRIDE2 1296, tile 0 the void tile. Element type is the
 * handle's precision: float for W2D_F32, double for W2D_F64. Static tables cross in GridTables' tile layout
 * (wadi2d/tables/builder.py:83-114), float32.
 */
#ifndef WADI2D_HIP_H
#define WADI2D_HIP_H

#include <stdint.h>

#include "wadi2d_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct w2d_handle w2d_handle;

/* GridTables (builder.py:83-114) plus TileGrid.nbr (grid/tiles.py:30-35); (ntiles, 32, 32[, rows[, 2]]). */
typedef struct w2d_tables {
    int32_t ntiles;
    int32_t rows;
    const int32_t* nbr; /* (ntiles, 8) neighbour tile ids in TileGrid.NEIGHBOURS order */
    const float* hc;
    const float* dtab;
    const float *flx, *frx, *glx, *grx, *efx, *cwx, *ftabx;
    const float *fly, *fry, *gly, *gry, *efy, *cwy, *ftaby;
    const float *zcx, *zcy;
} w2d_tables;

/* NumericsRecord (wadi2d/solver/backend.py:14-38); the kernels' Params (gpu/kernels/common.cuh:5-18). */
typedef struct w2d_params {
    double g, dx, e_wet, h_cfl, u_max, lead, a_min;
    int32_t second, sheet, mirror, carry;
} w2d_params;

enum { W2D_F32 = 4, W2D_F64 = 8 };
/* W2D_EXACT: IEEE division, square root and cube root, no contraction: the float32 statement order of
 * solver/cpu/kernels.py. W2D_FAST: hardware reciprocal, square root and log/exp cube root in float32. */
enum { W2D_EXACT = 0, W2D_FAST = 1 };
/* Or-ed into the variant: face-table lookups gather from the table every time instead of through the per-face
 * row cache (the results are the same; only the memory traffic differs). */
enum { W2D_NO_FACE_CACHE = 4 };

/* Field ids: layout.STATE_ARRAYS then layout.WORK_ARRAYS (layout.py:25-29). */
enum {
    W2D_H, W2D_D, W2D_U, W2D_V, W2D_MX, W2D_MY, W2D_QX, W2D_QY,
    W2D_U1, W2D_V1, W2D_MX1, W2D_MY1, W2D_QX1, W2D_QY1, W2D_R, W2D_SRC, W2D_RES, W2D_OM,
    W2D_NFIELDS
};

/* Backend constructor (GpuBackend.__init__, gpu/backend.py:58-104): uploads every table. */
int w2d_create(const w2d_tables* tables, const w2d_params* params, int32_t precision, int32_t variant,
               w2d_handle** out);
int w2d_destroy(w2d_handle* h);

/* set_field / get_field (gpu/backend.py:121-125): n = ntiles * 1296 entries of the handle's type. get_field fills
 * every halo entry from the tile that owns the cell, as the reference's exchange leaves them. The face momentum
 * (MX, MY) is not stored by a step: the face kernel recomputes the last step's 0.5 (hL + hR) dx u from the velocity
 * and the depths that step read, bit for bit the product it replaces; get_field and read_cells of MX, MY store it
 * first, and a set_field of H, U, V, MX or MY stores it before the change, so the next step takes the momentum as
 * the reference would hold it. MX1 and MY1 read and write MX and MY (one momentum, both buffers). */
int w2d_set_field(w2d_handle* h, int32_t field, const void* host, int64_t n);
int w2d_get_field(w2d_handle* h, int32_t field, void* host, int64_t n);
/* Model.read_cells (model.py:139-143): entries at own-cell state indices. */
int w2d_read_cells(w2d_handle* h, int32_t field, const int64_t* index, int64_t n, void* out);
/* No counterpart in the reference: w2d_read_cells that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The index list and the
 * gathered values go through a file-scope device array of 262,144 entries (4 MiB + 256 B) that each device holds
 * until the process ends; more cells in one call are refused (status 1). Every call on one device shares that array:
 * calls that overlap on one device, from several threads or handles, race. Same results as w2d_read_cells. */
int w2d_read_cells_noalloc(w2d_handle* h, int32_t field, const int64_t* index, int64_t n, void* out);
/* The writes the reference's callers make through its arrays (momentum.py:49-54 FaceKick.apply into mx; tests into u,
 * v, h): values (the handle's type) into own cells at state indices, under set_field's rules (H both buffers, the
 * momentum stored first and written to both buffers); an index given twice keeps its last value. */
int w2d_write_cells(w2d_handle* h, int32_t field, const int64_t* index, int64_t n, const void* values);

/* set_active (gpu/backend.py:116-119). */
int w2d_set_active(w2d_handle* h, const int32_t* tiles, int32_t n);
/* set_boundary (gpu/backend.py:127-132), BoundaryCells (solver/backend.py:41-55) at state indices; a cell's ref is read only
 * for kind 2 (it follows that cell), so kinds 0 and 1 may carry any index, as the reference's callers pass. */
int w2d_set_boundary(w2d_handle* h, const int32_t* cell, const int32_t* ref, const int32_t* kind, const float* d0,
                     const float* h0, int32_t n);
/* step (gpu/backend.py:151-188): faces, positivity, update, boundary; then the face buffers swap. */
int w2d_step(w2d_handle* h, double dt, int32_t want_wave, int32_t want_wet);
/* max_wave (gpu/backend.py:196-197). */
int w2d_max_wave(w2d_handle* h, double* out);
/* The per-tile wave speeds of the last wave step (the backend's `wave`, ntiles entries). */
int w2d_wave(w2d_handle* h, double* out);
/* wet_flags (gpu/backend.py:199-200) and the host's write of them (model.py:53). */
int w2d_wet_flags(w2d_handle* h, int32_t* out);
int w2d_set_wet(w2d_handle* h, const int32_t* flags);
/* refresh_plan / refresh_halos (gpu/backend.py:139-149). No halos are stored on the device, so the second has
 * nothing to do; it is kept for the protocol. */
int w2d_refresh_plan(w2d_handle* h, const int32_t* tiles, int32_t n);
int w2d_refresh_halos(w2d_handle* h);
/* Tally (engine/maxima.py:149-192): FP64 volume through a face list, every face its own running total. */
int w2d_tally_set(w2d_handle* h, const int32_t* face, const int32_t* axis, const float* sign, int32_t n);
int w2d_tally_add(w2d_handle* h, double dt);
int w2d_tally_volume(w2d_handle* h, double* out);
/* Every face's own running total (n as set), for books per ensemble member. */
int w2d_tally_totals(w2d_handle* h, double* out);
/* Puts saved per-face totals back, n the tally's face count (engine/maxima.py:189-192 Tally.restore, for a
 * checkpoint's resume, io/checkpoint.py:73). */
int w2d_tally_restore(w2d_handle* h, const double* totals, int32_t n);
/* The instantaneous flow through the tally's faces, sum of sign * q, m3/s (no counterpart: the outfall discharge). */
int w2d_tally_flow(w2d_handle* h, double* out);
/* Maxima (engine/maxima.py:28-51,103-146): depth, mean depth, speed, hazard, arrival; which = 0..4. */
int w2d_maxima_set(w2d_handle* h, double arrival_depth, double speed_depth);
int w2d_maxima_update(w2d_handle* h, double t);
int w2d_maxima_get(w2d_handle* h, int32_t which, void* host, int64_t n);
/* Puts saved maxima back, state layout as maxima_get gives them (engine/maxima.py:136-138: Maxima.restore, for a
 * checkpoint's resume, io/checkpoint.py:59-75). */
int w2d_maxima_put(w2d_handle* h, int32_t which, const void* host, int64_t n);
/* Monitor.storage (engine/monitor.py:214-218) without the dx^2: sum of h + res over listed own cells, FP64. */
int w2d_storage_set(w2d_handle* h, const int64_t* index, int64_t n);
int w2d_storage(w2d_handle* h, double* out);
/* A source that is a fixed weight per cell times one number per window (no counterpart: the reference uploads the
 * window's float32 rates, model.py:112-122): weights at own-cell state indices, then per window
 * src = float(weight * scale) on the device, and the sum of src over all cells. */
int w2d_source_weights(w2d_handle* h, const int64_t* index, const double* weight, int64_t n);
int w2d_source_scale(w2d_handle* h, double scale);
int w2d_source_sum(w2d_handle* h, double* out);
/* The source as the rain excess of each cell (forcing.py:644-678 with hydrology/losses.py:49-77): per own cell a
 * depth factor w (storm.depth_grid's depth, forcing.py:594-601; 1 without one) and a curve number cn (NaN: no losses,
 * forcing.py:186-203). A window [t0, t1] over which the storm's cumulative depth runs from d0 to d1 mm gives each
 * cell float((E(d1 w) - E(d0 w)) / denom), denom = 1000 (t1 - t0), where E(P) = P without a curve number, else the
 * SCS cumulative excess max(P - Ia, 0)^2 / (max(P - Ia, 0) + S), S = (1000 / cn - 10) 25.4 mm, Ia = 0.2 S, NaN
 * read as 0. Unlisted cells rate 0. The form holds until w2d_source_weights is called. */
int w2d_source_losses(w2d_handle* h, const int64_t* index, const double* w, const double* cn, int64_t n);
/* One window's rates into every own cell's src; total = their FP64 sum (in tile order), and tiles[t] (ntiles, or
 * null) = 1 where a tile has a non-zero rate: the tiles set_source_cells keeps active (model.py:112-122). */
int w2d_source_excess(w2d_handle* h, double d0, double d1, double denom, double* total, int32_t* tiles);
/* Gridded rain (storm/gridded.py, hydrology/cell_rain.py:19-58): after w2d_source_losses, each listed own cell's
 * pixel in [0, npix) (the cell reads the storm depths of that pixel instead of d0 and d1; unlisted cells take none).
 * The form holds until w2d_source_losses or w2d_source_weights is called. */
int w2d_source_pixels(w2d_handle* h, const int64_t* index, const int32_t* pixel, int64_t n, int32_t npix);
/* w2d_source_excess for gridded rain: the npix pixels' cumulative depths at the window's start (d0) and end (d1). */
int w2d_source_excess_pixels(w2d_handle* h, const double* d0, const double* d1, int32_t npix, double denom,
                             double* total, int32_t* tiles);
/* Kernel timing (no counterpart): with timing on, each step records HIP events on the handle's stream around its
 * face and cell kernels; timing_read waits, and returns and clears the summed milliseconds and the step count. */
int w2d_timing(w2d_handle* h, int32_t on);
int w2d_timing_read(w2d_handle* h, double* faces_ms, double* cells_ms, int64_t* launches);
/* A plain device copy of `bytes` (no counterpart): the yardstick that scales the memory counters; returns the mean
 * milliseconds per copy over `repeats` copies. */
int w2d_copy_probe(int64_t bytes, int32_t repeats, double* ms);
/* Device-resident run (no counterpart: it is the loop of Stepper.advance, solver/stepper.py:101-129, and of the
 * run's source windows, run.py:1712-1739, moved onto the device, for one model or an ensemble). Members own
 * contiguous tile ranges of the handle; each keeps its own clock, step length, wave refresh, source windows, rain
 * and active tiles. Window k of a member spans [t0 + k window, min(t0 + (k+1) window, t_end)]; scales[offset + k] is
 * its source rate per unit weight (see w2d_source_weights), m/s. */
typedef struct w2d_member {
    int32_t tile0, tile1;
    double t0, t_end, dt0, window;
    int32_t nwindows;
    int64_t scale_offset;
    int64_t nsteps0;
    int32_t refresh_due;
    int32_t dry;  /* the last window's rain was nothing (run.py:1708, 1729): a resumed run's state */
} w2d_member;
/* maxima: the running maxima are sampled after every step whose count is a multiple of it (0: not sampled), as
 * the run loop's maxima_every (engine/maxima.py:103-146, config/schema.py:1001). */
typedef struct w2d_run_numerics {
    double cfl, dt_max, dt_min, wave_floor;
    int32_t dt_every, active_every, maxima;
} w2d_run_numerics;
typedef struct w2d_member_state {
    double t, dt, rain;
    int64_t nsteps;
    int32_t done, failed, refresh_due, dry;
    double dt_min; /* the smallest step length dt held after any step of this run (the run summary's dt_min) */
    double q_peak, t_peak; /* the largest discharge out through the member's tally faces after a step of this run
                            * (the FP64 sum of sign * q), and the time it was reached */
    int32_t pipelined;     /* the member takes the pipelined step (w2d_run_pipeline) */
    int64_t pipelined_updates; /* cell updates of this run made by the pipe kernel (the rest by the cell kernel) */
} w2d_member_state;
/* With the excess source (w2d_source_losses), a device-resident run takes (d0, d1, denom) per window, in the order of
 * the scales (3 nwindows doubles, nwindows = nscales), before w2d_run_setup: its windows' rates are formed on the
 * device as w2d_source_excess forms them, a window opens when its excess falls or the last one's did, and its kept
 * tiles are those with a non-zero rate (run.py:1719-1729). */
int w2d_run_terms(w2d_handle* h, const double* terms, int64_t nwindows);
/* With gridded rain, the pixels' depths of every window as well, before w2d_run_setup: nwindows rows of npix depths
 * at the window's start then npix at its end, in the order of the terms. */
int w2d_run_pixel_terms(w2d_handle* h, const double* depths, int64_t nwindows, int32_t npix);
/* Set up a run from the handle's current state, active list and the stepper's idle counts (ntiles or null). */
int w2d_run_setup(w2d_handle* h, const w2d_member* members, int32_t M, const double* scales, int64_t nscales,
                  const w2d_run_numerics* num, const int32_t* idle);
/* Enqueue nsteps steps of every live member; with graph_steps > 0 they are replayed from one captured graph of that
 * many steps (not while kernel timing is on: HIP does not time events inside graphs). Nothing waits. Each member keeps
 * its own buffer parity on the device (the steps it took), so a member that is held or done while others step keeps
 * its state where it is, and one graph serves every parity. */
int w2d_run_steps(w2d_handle* h, int32_t nsteps, int32_t graph_steps);
/* Wait, and read every member's clock, step, count, rain volume and flags (failed: 1 non-finite wave, 2 step below
 * the minimum). Like every call that reads or writes the state after run_steps (get_field, read_cells, set_field,
 * write_cells, set_active, step, tally_add, storage, maxima_update, refresh_plan), it first applies the list check the
 * last step asked for and brings every member to one buffer parity, so the fields read are each member's own. */
int w2d_run_read(w2d_handle* h, w2d_member_state* out, int32_t M);
/* Hold each member once its clock t >= t_stop[m] (no counterpart: the host reads outputs at the step that reaches
 * an output time, as the host loop's step hook does); +inf lets it run on. Steps enqueued while a member is held
 * change nothing of it, whatever the other members do. */
int w2d_run_pause(w2d_handle* h, const double* t_stop, int32_t M);
/* The pipelined step of device-resident runs (no counterpart; on by default in FP32, off in FP64; taken by the run
 * set up, from its next step, and by the next w2d_run_setup): a step whose next step keeps the active list and reads
 * no wet flag (no check or window end follows, or one that cannot change the list: every tile listed, and for a
 * check every tile kept whatever its wet flag) leaves its cell update to that step's kernel, which makes it and then
 * the face step in one pass over each tile, recomputing the update of its west column and north row instead of
 * reading the level and limited discharges back from memory. Every field, maximum and book is
 * the two-kernel step's, bit for bit; run_steps returns with no update waiting. Runs with the learned closure, and
 * members with a following boundary cell (kind 2) whose reference lies outside what a tile updates, take the
 * two-kernel step. on: 0 or 1. */
int w2d_run_pipeline(w2d_handle* h, int32_t on);
int w2d_run_idle(w2d_handle* h, int32_t* out);
/* The active list as the device last left it (out may be null to ask for the count). */
int w2d_active_tiles(w2d_handle* h, int32_t* out, int32_t* n);
/* The learned face closure (no counterpart in the reference, which has none): n faces, each the west (axis 1) or
 * north (axis 0) face of own-layout cell cell_r[f] (t * 1024 + j * 32 + i) with cell_l[f] across it; stat[f][11]
 * their static features; norm the 16 feature means, 16 inverse deviations and the scale on the invert-rise output;
 * params the network's FP32
 * parameters (w2d_mlp_*). Before every step that refreshes the wave speed (and so before the first) the network
 * corrects those faces from the current levels: their inverts rise and, with `conveyance`, their K^(3/5) scales.
 * FP32 handles only; n = 0 turns it off. */
int w2d_closure_set(w2d_handle* h, int32_t n, const int32_t* cell_r, const int32_t* cell_l, const int32_t* axis,
                    const float* stat, const float* norm, const float* params, int32_t conveyance);
/* Workgroup sizes of the face and cell kernels, 256 or 512 threads each (defaults 512, 256); the fields do not
 * depend on them. No counterpart in the reference, whose launch shape is fixed (gpu/backend.py:151-188): the
 * launch-shape tuner's knob. */
int w2d_set_launch(w2d_handle* h, int32_t faces_threads, int32_t cells_threads);
/* synchronize (gpu/backend.py:202-203). */
int w2d_synchronize(w2d_handle* h);

/* ---- Local time stepping: the operations of the level stepper, one call per method of GpuLevels
 * (solver/gpu/levels.py:11-203, binding solver/gpu/kernels/levels.cu:1-172 and the tile-list faces, limit and
 * update). A level run keeps the handle's buffer parity: the depth is updated in place, the face velocities and
 * momenta step into the other buffers and are committed tile by tile. Tiles hold no halos, so the reference's
 * halo exchanges (faces_done, limit_done) have no call; halo_state only sets the smallest step at which the staged
 * neighbours are read, between a coarser neighbour's old and new depth and level.
 * setup: allocates the ratio, old depth/level and edge estimates, and stores the face momentum. */
int w2d_level_setup(w2d_handle* h);
/* set_level (gpu/levels.py:61-62): every tile's time level. */
int w2d_level_set(w2d_handle* h, const int32_t* level);
/* lists (gpu/levels.py:58-59): a list of tiles (width 1) or of edge rows (width 4: fine tile, side of the fine tile
 * 0 W 1 N 2 E 3 S, coarse tile, edge number < 4 ntiles), copied to the device; `id` names it until lists_clear. */
int w2d_level_list(w2d_handle* h, const int32_t* a, int32_t n, int32_t width, int32_t* id);
int w2d_level_lists_clear(w2d_handle* h);
/* halo_state (gpu/levels.py:64-80, halo_levels levels.cu:67-91): the smallest step `sub` of the cycle. */
int w2d_level_halo_state(w2d_handle* h, int32_t sub);
/* faces (gpu/levels.py:96-103; faces.cu:244-271 over a list at one dt). */
int w2d_level_faces(w2d_handle* h, int32_t tiles, double dt, int32_t want_wave);
/* strips (gpu/levels.py:105-106; levels.cu:34-64): edge faces the coarse tile owns, at the fine rate. */
int w2d_level_strips(w2d_handle* h, int32_t rows, double dt, int32_t want_wave);
/* limit (gpu/levels.py:146-150; cells.cu:4-18 over a list at one dt). */
int w2d_level_limit(w2d_handle* h, int32_t tiles, double dt);
/* save (gpu/levels.py:155-157; save_old levels.cu:93-105). */
int w2d_level_save(w2d_handle* h, int32_t tiles);
/* estimate (gpu/levels.py:159-161; edge_estimate levels.cu:143-154). */
int w2d_level_estimate(w2d_handle* h, int32_t rows);
/* flux (gpu/levels.py:163-168; edge_flux levels.cu:156-172): one axis's rows per call. */
int w2d_level_flux(w2d_handle* h, int32_t rows, double dt);
/* update (gpu/levels.py:175-181; cells.cu:28-63 with debt): every listed tile's wet flag is written. */
int w2d_level_update(w2d_handle* h, int32_t tiles, double dt, int32_t want_plan);
/* commit and commit_strips (gpu/levels.py:189-199; levels.cu:107-140). */
int w2d_level_commit(w2d_handle* h, int32_t tiles);
int w2d_level_commit_strips(w2d_handle* h, int32_t rows);
/* clear_wave (what 0) and clear_wet (what 1) (gpu/levels.py:86-90). */
int w2d_level_clear(w2d_handle* h, int32_t what);
/* boundary (gpu/levels.py:183-187; cells.cu boundary) on the in-place depth. */
int w2d_level_boundary(w2d_handle* h);
const char* w2d_last_error(void);

/* ---- The learned face closure's network (16 -> 64 -> 64 -> 2, ReLU; BF16 weights on matrix cores, FP32
 * accumulation). No counterpart in the reference, which has no learned component: these serve the closure's
 * training driver. Pointers are device pointers; `stream` is a hipStream_t (null: the default stream); nothing
 * waits. Parameter layouts: FP32 vector W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]; BF16 block W1,
 * W2, W3 padded to [16][64]. Inputs x are BF16 [n][16], outputs FP32 [n][2]. */
int w2d_mlp_sizes(int32_t* nparam, int32_t* nbf16);
/* Plain device memory for the training driver's arrays: allocate, free, copy (kind 0 host to device, 1 device to
 * host, 2 device to device; synchronous), wait for the device. */
int w2d_dev_alloc(int64_t bytes, void** out);
int w2d_dev_free(void* p);
int w2d_dev_copy(void* dst, const void* src, int64_t bytes, int32_t kind);
int w2d_dev_synchronize(void);
/* A minibatch: rows idx[0 .. n) of x [.][16] (BF16), target and mask [.][2], into dense arrays. */
int w2d_mlp_take(const uint16_t* x, const float* target, const float* mask, const int64_t* idx, int64_t n,
                 uint16_t* x_out, float* target_out, float* mask_out, void* stream);
/* BF16 weight block from the FP32 parameters. */
int w2d_mlp_pack(const float* params, uint16_t* wbf16, void* stream);
int w2d_mlp_forward(const uint16_t* wbf16, const float* params, const uint16_t* x, int64_t n, float* y, void* stream);
/* Workgroups w2d_mlp_gradient uses for n samples (its partial buffer holds blocks * (nparam + 1) floats). */
int w2d_mlp_train_blocks(int64_t n, int32_t* blocks);
/* Gradient of sum over samples and outputs of mask * (y - target)^2 / norm with respect to the FP32 parameters
 * (through the BF16 forward pass), into grad[0 .. nparam), and the loss into grad[nparam]. */
int w2d_mlp_gradient(const uint16_t* wbf16, const float* params, const uint16_t* x, const float* target,
                     const float* mask, int64_t n, float norm, float* partial, int32_t blocks, float* grad,
                     void* stream);
/* One Adam step (step counts from 1) on the FP32 parameters, then their BF16 block. */
int w2d_mlp_adam(float* params, float* m, float* v, const float* grad, int32_t step, float lr, float beta1,
                 float beta2, float eps, uint16_t* wbf16, void* stream);
/* Devices the HIP runtime sees (select_backend's probe, solver/backend.py:58-75). */
int w2d_device_count(int32_t* out);

/* build_tables (tables/builder.py:131-258) on the device: the same arguments, results and statements as
 * w2d_build_tables (wadi2d_host.h), one device thread per cell in float64, the host's arrays copied in and the
 * results copied back before it returns. The reference builds its tables on the CPU only (builder.py:147-148
 * spreads tiles over processes); this is the faster builder for large grids. The device's cbrt and pow may differ
 * from the host library's in the last bit, so a row may differ from the host builder's by one float32 ulp.
 * Returns 0, 1 on a bad argument, 2 when the device is out of memory, 3 on another HIP error. */
int w2d_build_tables_device(const double* bed, int64_t bed_ny, int64_t bed_nx, int64_t row_off, int64_t col_off,
                            double pixel, int32_t ny, int32_t nx, double dx, const int32_t* tiles, int32_t ntiles,
                            const uint8_t* mask, const double* manning, int32_t nman_grid, const double* crest_xz,
                            const double* crest_xlen, const double* crest_xc, const double* crest_yz,
                            const double* crest_ylen, const double* crest_yc, const w2d_table_config* cfg,
                            w2d_table_out* out);

#ifdef __cplusplus
}
#endif

#endif
