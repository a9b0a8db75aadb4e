/* gpu_phs for MI355X — C ABI of libgpu_phs.so.
 *
 * Two groups of entry points:
 *
 * 1. The Whizard-facing interface, the same 14 extern "C" symbols with the same signatures,
 *    argument meaning and index conventions as the reference's whizard_gpu_phs.cpp (each
 *    declaration cites the function it replaces).  Arguments are passed by pointer, channel
 *    numbers are 1-based, trees use bincodes, output buffers are caller-owned host arrays.
 *    The reference returns nothing and reports nothing; this library keeps the void
 *    signatures and records the status of every call for gpu_phs_last_error().
 *
 * 2. Additions with no counterpart in the reference: process handles, device-resident
 *    batched plans (one launch, or one hipGraph, over many channels and processes),
 *    per-segment plain launches for comparison, an autotuner, and the adjoint (gradients
 *    with respect to the randoms and the map masses and widths).  These return a status
 *    (GPU_PHS_OK = 0, negative on error) or a non-negative handle.
 *
 * Layouts (both groups): x [n_events][n_x] event-major, n_x = 3 n_out - 4, randoms consumed
 * as msq randoms of the internal non-root branches in post-order, then (phi, cos theta) of
 * every internal branch in pre-order; p [n_events][n_out][4] with final particle j at slot j
 * (bincode 2^j); factor, volume [n_events]; ok [n_events] one byte (C++ bool / c_bool).
 */
#ifndef GPU_PHS_H
#define GPU_PHS_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ---------------------------------------------------------------------------- */
#define GPU_PHS_OK 0
#define GPU_PHS_E_ARG (-1)        /* invalid argument or inconsistent sizes */
#define GPU_PHS_E_TREE (-2)       /* invalid tree or unsupported map id */
#define GPU_PHS_E_STATE (-3)      /* call out of order (e.g. generate before init) */
#define GPU_PHS_E_HIP (-4)        /* HIP runtime error */
#define GPU_PHS_E_HANDLE (-5)     /* unknown process or plan handle */
#define GPU_PHS_E_IO (-6)         /* file could not be read or parsed */
#define GPU_PHS_E_NODEVICE (-7)   /* no GPU visible */

/* Status of the last call on this thread (0 = success) and its message. */
int gpu_phs_last_error(void);
const char* gpu_phs_last_error_message(void);

/* ---- 1. Whizard interface (whizard_gpu_phs.cpp) ---------------------------------------- */

/* whizard_gpu_phs.cpp:12-33: sizes of the process; allocates the flavour tables. */
void c_whizard_set_particle_structure(int* n_channels, int* n_trees, int* n_groves, int* n_x, int* n_in,
                                      int* n_out);
/* whizard_gpu_phs.cpp:35-40: masses and widths of the n_in + n_out externals, incoming first. */
void c_whizard_set_flavors(double* masses, double* widths);
/* whizard_gpu_phs.cpp:42-55: allocates the per-channel mapping tables. */
void c_whizard_init_mappings(void);
/* whizard_gpu_phs.cpp:57-64: map ids, masses, widths of one channel (1-based), each
 * n_prt_out = 2^n_out - 1 long, index = bincode - 1. */
void c_whizard_fill_mapping(int* channel, int* map_ids, double* masses, double* widths);
/* whizard_gpu_phs.cpp:66-77: allocates the per-channel tree tables. */
void c_whizard_init_tree_structures(void);
/* whizard_gpu_phs.cpp:79-87: one channel's tree (1-based channel), each n_prt =
 * 2^(n_in+n_out) - 1 long: daughters as bincodes (0 = none), has_children flags, friends
 * (bincode of the incoming particle whose beam a branch's polar angle is measured from,
 * read as phs.cu:276-284 reads it: 2^n_out -> the beam along -z, 2^(n_out+1) -> the beam
 * along +z, any other value -> the branch's own z axis).  The rotation onto the beam is a
 * proper rotation (Rodrigues), where the reference's L_rot_to_2nd (phs.cu:753-781) is not
 * orthogonal. */
void c_whizard_fill_tree_structure(int* channel, int* daughters1, int* daughters2, int* has_children,
                                   int* friends);
/* whizard_gpu_phs.cpp:89-98: channel of every event of the next batch from cumulative
 * per-channel end indices.  Unlike the reference, empty channels anywhere in the list are
 * skipped correctly (the reference only skips leading ones). */
void c_whizard_init_channel_ids(int* batch_size, int* n_channels, int* channel_limits);
/* whizard_gpu_phs.cpp:100-104: thread counts of the reference's three kernels.  Here the
 * first value (rounded to a multiple of 64, at most 256) sets the workgroup size of the fused
 * kernel; the others are accepted and ignored. */
void c_whizard_set_threads(int* msq_threads, int* cb_threads, int* ab_threads);
/* whizard_gpu_phs.cpp:106-108: compiles the trees into kernel programs and uploads them. */
void c_whizard_init_gpu_phs(void);
/* whizard_gpu_phs.cpp:110-112: releases everything the calls above allocated. */
void c_whizard_finalize_gpu_phs(void);
/* whizard_gpu_phs.cpp:114-122: generates n_events events from host randoms x (event-major,
 * n_x each) with per-event beam energies E1, E2 (equal for all events, or differing: then
 * each event gets its own sqrt(s), phs.cu:1313-1317).  Writes p [n_events][n_out][4],
 * factors at [n_channels * i + c] for every channel c (the event's own channel: its factor;
 * the others: the inverse mapping of its momenta, phs.cu:1491-1554), volumes [n_events], oks
 * [n_events], and x_out [n_channels][n_events][n_x] (own channel: the input randoms; the
 * others: the inverse mapping; may be NULL).  requires_boost_to_lab is read but, as in the
 * reference (phs.cu:1559), the lab boost is applied whenever E1 != E2.  The host copies go
 * through pinned buffers, each chunk split over the calling thread and 7 worker threads that
 * the library starts at the first large call and keeps until it is unloaded. */
void c_whizard_gen_phs_from_x_gpu(double* E1_in, double* E2_in, bool* requires_boost_to_lab, int* n_events,
                                  int* n_channels, int* n_x, double* x, double* factors, double* volumes,
                                  bool* oks, double* p, double* x_out);
/* whizard_gpu_phs.cpp:124-126: the reference leaves this empty.  Here it writes, into the
 * first 8 bytes of p, the device address of the momenta of the last generate call
 * (layout [n_events][n_out][4]), or 0 if there is none. */
void c_whizard_get_momentum_device_pointer(double* p);
/* whizard_gpu_phs.cpp:128-130 (reset_phs, phs.cu:1303-1308): drops the flavour tables and
 * the device buffers kept between calls. */
void c_whizard_reset_generator(void);
/* whizard_gpu_phs.cpp:132-249: prints the module state (sizes, maps, trees, programs). */
void c_whizard_show_module(void);

/* ---- 2. Additions ---------------------------------------------------------------------- */

/* A process: n_channels trees in the reference format (see c_whizard_fill_tree_structure
 * and c_whizard_fill_mapping; arrays are [n_channels][n_prt] and [n_channels][n_prt_out]),
 * flavour masses [n_in + n_out], beams E1, E2 (sqrt(s) = 2 sqrt(E1 E2)).  friends may be
 * NULL (no branch has a friend; semantics as in c_whizard_fill_tree_structure).  Returns a
 * handle >= 0. */
int gpu_phs_process_create(int n_in, int n_out, int n_channels, const int* d1, const int* d2,
                           const int* has_children, const int* friends, const int* map_ids,
                           const double* map_masses, const double* map_widths, const double* flv_masses,
                           const double* flv_widths, double E1, double E2);
/* A process from a reference-format .ref file (file_input.cpp:123-210 layout: header,
 * trees, flavours; events are not read). */
int gpu_phs_process_load_ref(const char* path, double E1, double E2);
int gpu_phs_process_destroy(int process);
/* n_out, n_x, n_channels of a process. */
int gpu_phs_process_info(int process, int* n_out, int* n_x, int* n_channels);
/* The compile-time tree shape each channel's forward kernel is specialised for (an index of
 * the library's shape table: the ordered trees of 2, 3 or 4 final-state particles), or -1 where
 * the generic kernel runs it (larger trees, friends).  n >= n_channels. */
int gpu_phs_process_shapes(int process, int* shapes, int n);
/* The process's tree arrays as created (sizes from gpu_phs_process_info); any pointer may
 * be NULL. */
int gpu_phs_process_trees(int process, int* d1, int* d2, int* has_children, int* map_ids,
                          double* map_masses, double* map_widths, double* flv_masses);

/* One contiguous run of events of one channel of one process, with device buffers. */
typedef struct {
  int process;            /* handle */
  int channel;            /* 0-based */
  int64_t n_events;
  const double* x;        /* device [n_events][n_x] */
  double* p;              /* device [n_events][n_out][4] */
  double* factors;        /* device [n_events] */
  double* volumes;        /* device [n_events] */
  uint8_t* oks;           /* device [n_events] */
} gpu_phs_segment;

/* A plan: the tile table of all segments, uploaded once.  Returns a handle >= 0. */
int gpu_phs_plan_create(const gpu_phs_segment* segments, int n_segments);
int gpu_phs_plan_destroy(int plan);
/* Launch shape: workgroup size (64..256, multiple of 64), events per thread per tile
 * (1..16), persistent grid (0 = one workgroup per tile), tile order (0 = segment order,
 * 1 = tiles of the deepest trees first), kernel variant: waves_per_eu = register budget for
 * at least that many waves per SIMD (0 = compiler's choice, 3, 4, or 5: shape-specialised
 * kernels only, others take 5 as 0), lds = bit flags: 1 keeps
 * the boosts of deferred subtrees in LDS instead of VGPRs (trees that defer two levels always
 * use LDS), 2 keeps each event's randoms in LDS instead of VGPRs (always for n_x > 16).  ept 1
 * with grid 0 (one workgroup per tile of one sub-tile) runs the shape-specialised kernels'
 * one-pass variants at the compiler's register budget (fewer registers: four waves per SIMD for
 * trees of up to four final-state particles, three for five and six).  Changing block, ept or order
 * rebuilds the tile table.  A block whose LDS tables (4 n_out rows of block + 1 doubles, n_x
 * more with randoms in LDS, plus the LDS save banks) exceed the device's per-workgroup LDS is
 * rejected. */
int gpu_phs_plan_set_config(int plan, int block, int ept, int grid, int order, int waves_per_eu, int lds);
int gpu_phs_plan_get_config(int plan, int* block, int* ept, int* grid, int* order, int* waves_per_eu,
                            int* lds);
/* Launch shape of one tile class (0: trees of up to four final-state particles whose only
 * deferred subtree is the root's; 1: the others), arguments as gpu_phs_plan_set_config, which
 * sets both classes alike.  The autotuner sets them separately; these reproduce its choice. */
int gpu_phs_plan_set_class_config(int plan, int cls, int block, int ept, int grid, int waves_per_eu, int lds);
int gpu_phs_plan_get_class_config(int plan, int cls, int* block, int* ept, int* grid, int* waves_per_eu, int* lds);
/* Tile order (as in gpu_phs_plan_set_config) and whether the two tile classes run one after the
 * other on the caller's stream (serial = 1) or share the GPU (0: the second class on a side
 * stream, joined back by events). */
int gpu_phs_plan_set_schedule(int plan, int order, int serial);
/* Tree shapes with fewer than `events` events in a tile class run in that class's one launch of
 * the generic kernel instead of a launch of the kernel specialised for them (default 2^18; 0:
 * every shape its own launch).  A launch however small lasts about one wave's lifetime, so
 * small shapes are cheaper folded into a launch that runs anyway.  No counterpart in the
 * reference. */
int gpu_phs_plan_set_merge_below(int plan, int64_t events);
/* Number of streams (1..4) the plan's launches spread over: with n_streams > 1 every launch of
 * the plan (one per tree shape, plus the generic one) goes to one of the caller's stream and
 * n_streams - 1 side streams, heaviest first onto the least loaded, joined back into the
 * caller's stream by events; the classes' serial flag then does not apply.  A launch ends with
 * its slowest workgroups while the others' slots idle; launches on several streams fill those
 * slots.  Default 1.  No counterpart in the reference. */
int gpu_phs_plan_set_streams(int plan, int n_streams);
/* The device's streaming copy bandwidth in GB/s (bytes read plus bytes written per second,
 * median of `reps` copies of `bytes` bytes with a kernel of one 16-byte word per thread and
 * streaming stores): a practical ceiling for the memory-bound kernels, next to the published
 * peak.  Allocates two buffers of `bytes` bytes for the call.  No counterpart in the
 * reference. */
int gpu_phs_copy_bandwidth(int64_t bytes, int reps, void* stream, double* gb_s);
int gpu_phs_plan_get_streams(int plan, int* n_streams);
int gpu_phs_plan_get_merge_below(int plan, int64_t* events);
int gpu_phs_plan_get_schedule(int plan, int* order, int* serial);
/* Per-event beam energies of segment `segment` (device arrays [n_events], or both NULL for the
 * process's fixed beams): each event gets sqrt(s) = 2 sqrt(E1 E2), its lab boost and the map
 * constants that depend on sqrt(s), as the reference does per event (phs.cu:1313-1317,
 * 1286-1301; mappings_gpu.hpp).  The arrays must outlive the plan's use.  The adjoint of a
 * plan with per-event beams takes external cotangents only (loss_mode 0). */
int gpu_phs_plan_set_event_beams(int plan, int segment, const double* E1, const double* E2);
/* Per-branch msq output of segment `segment`: a device array [n_events][2^n_out - 1] that
 * receives, at index bincode - 1, the invariant mass squared of every branch of the event's
 * tree -- m_j^2 for the final-state particles, s for the root, the mapped msq of the others --
 * as the reference's per-branch array holds them after its decay stage (phs.cu:570-624,
 * phs_cpu.cpp:174-234).  Slots of bincodes that are not branches of the tree are not written.
 * NULL (the default) turns it off; it costs no memory traffic then.  The array must outlive
 * the plan's use. */
int gpu_phs_plan_set_msq_output(int plan, int segment, double* msq);
/* 1 (the default): channels without friends run forward kernels specialised for their tree
 * shape (every ordered tree of 2-6 final-state particles; with per-event beam energies, those of
 * 2-4); 0: every channel runs the generic kernel.  Both give identical results. */
int gpu_phs_plan_set_specialize(int plan, int on);
/* Forward pass of every segment in one kernel launch on `stream` (hipStream_t, 0 = null). */
int gpu_phs_plan_run(int plan, void* stream);
/* Forward pass with one launch per segment (the plain per-channel form). */
int gpu_phs_plan_run_plain(int plan, void* stream);
/* Capture the single launch into a hipGraph (once) and replay it on `stream`. */
int gpu_phs_plan_run_graph(int plan, void* stream);
/* Try launch shapes, time each with hipEvents (median of `reps`), keep the fastest.  Writes
 * a one-line summary to report (may be NULL).  Outputs are overwritten by the trial runs. */
int gpu_phs_plan_autotune(int plan, void* stream, int reps, char* report, int report_len);
/* Number of tiles and segments of the plan, and the save-bank depth its trees need. */
int gpu_phs_plan_info(int plan, int* n_tiles, int* n_segments, int* n_banks);

/* Adjoint buffers of one segment (same order as the plan's segments). */
typedef struct {
  const double* g_p;      /* device [n_events][n_out][4] cotangents of the momenta, or NULL */
  const double* g_f;      /* device [n_events] cotangents of the factor, or NULL */
  double* g_x;            /* device [n_events][n_x]: d loss / d x (0 for events with ok = 0) */
  double* loss;           /* device [n_events]: per-event loss (loss_mode 1), or NULL */
} gpu_phs_adjoint_segment;

/* Reverse-mode derivatives of every segment in one launch.  loss_mode 0: the loss is
 * sum(g_p . p) + sum(g_f * factor) over ok events (a vector-Jacobian product); loss_mode 1:
 * the loss is sum over ok events of factor * sum_j (p_jx^2 + p_jy^2) / s, and g_p, g_f are
 * ignored.  Writes g_x, loss, the forward outputs when write_forward != 0, and gparam
 * [n_params]: d loss / d M at 2 slot, d loss / d Gamma at 2 slot + 1, per (channel, branch)
 * slot of gpu_phs_plan_param_map.  Covers friend trees and per-event beams (loss_mode 0
 * there, since the built-in loss assumes symmetric beams); trees whose tape (8 doubles per
 * branch, 19 per deferred level, 2 more per branch above four final-state particles, for each
 * of a workgroup's 64 events) exceeds the device's LDS per workgroup are rejected with
 * GPU_PHS_E_TREE: on MI355X, 2 -> 10 takes 75 KB of 160 KB.  The plan owns the partial sums and
 * per-channel completion counters of the reduction: adjoint calls on one plan must not overlap
 * (queue them on one stream). */
int gpu_phs_plan_adjoint(int plan, const gpu_phs_adjoint_segment* adjoint, int loss_mode, int write_forward,
                         double* gparam, void* stream);
/* 2 x the number of (channel, branch) parameter slots of a plan. */
int gpu_phs_plan_n_params(int plan);
/* For each slot: process handle, 0-based channel and bincode of the branch whose map mass and
 * width it refers to.  Arrays hold at least n_params / 2 entries. */
int gpu_phs_plan_param_map(int plan, int* process, int* channel, int* bincode, int n);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  The forward pass of
 * n_events events of one channel (0-based) of a process, as a one-segment plan would compute them
 * (same kernels, same results), with device buffers x [n_events][n_x], p [n_events][n_out][4],
 * factors, volumes [n_events] and oks [n_events], on `stream`.  It allocates nothing: the call's
 * tables go into a static 1 MiB device array that every GPU the library runs on holds until the
 * process ends, which limits a call to 64,512 tiles (66,060,288 events of a tree of up to four
 * final-state particles, 16,515,072 of a larger one); larger calls are rejected with
 * GPU_PHS_E_ARG.
 * All calls on one GPU share that array and overwrite it: calls that overlap on one GPU race, so
 * order them on one stream. */
int gpu_phs_generate_noalloc(int process, int channel, int64_t n_events, const double* x, double* p,
                             double* factors, double* volumes, uint8_t* oks, void* stream);

/* Inverse mapping (replaces the do_inverse_mapping branch of gen_phs_from_x_gpu,
 * phs.cu:1491-1554, and its kernels phs.cu:626-675, 924-1284): for every event of the
 * plan's segments and every channel c of their process, the randoms x_out and the factor
 * under which channel c would have produced the event's momenta.  Run after the forward
 * pass (it reads the segments' p, x and factor).  All segments must come from one process.
 * seg_row0[s] is the output row of segment s's first event; outputs are device arrays
 * factors [n_rows][n_channels] and x_out [n_channels][n_rows][n_x] (channel-major, as
 * Whizard reads them, phs.cu:222, 1585).  The generating channel's slot gets the forward
 * factor and randoms.  block: threads per workgroup (64..256; 0: 256 for processes of at most
 * 16 channels, 128 for more, smaller where the LDS tables do not fit).  Where the reference
 * has defects (the STEP_H inverse, the T/U inverse of the CPU version, the friend rotation)
 * the inverse of the generating channel returns its randoms instead. */
int gpu_phs_plan_inverse(int plan, const int64_t* seg_row0, int64_t n_rows, double* factors, double* x_out,
                         int block, void* stream);

/* Adaptive grids of the randoms (replaces rng.cpp:18-93: init_rng, rng_generate,
 * update_weights, which run on the host).  A grid holds, per channel of a process and per
 * random, 10 equal bins with weights, uniform at creation (rng.cpp:20-27).  Returns a handle
 * > 0. */
int gpu_phs_grid_create(int process);
int gpu_phs_grid_destroy(int grid);
/* The weights [n_channels][n_x][10] (n = their count). */
int gpu_phs_grid_get_weights(int grid, double* weights, int n);
int gpu_phs_grid_set_weights(int grid, const double* weights, int n);
/* Fills the randoms x of every segment of `plan` (all of the grid's process) on the device:
 * random k of event e of segment s is drawn from its channel's piecewise-constant density
 * (rng.cpp:70-90), with uniform numbers from Philox4x32-10 at counter (e, s, k / 2) and key
 * `seed` instead of the reference's std::default_random_engine, so a call is reproducible. */
int gpu_phs_grid_generate(int grid, int plan, uint64_t seed, void* stream);
/* Sets the weights of every channel that has accepted events (ok) in `plan` to the histogram
 * of their randoms, bin count / accepted count (rng.cpp:33-67); channels without accepted
 * events keep theirs (the reference divides by zero there).  Synchronizes with `stream`. */
int gpu_phs_grid_update(int grid, int plan, void* stream);

#ifdef __cplusplus
}

/* transfer_to_whizard.h:1-3, with C++ linkage as the reference declares them (defined in
 * phs.cu:1638-1644; get_ok_device_pointer is declared there but never defined): device
 * pointers of the last c_whizard_gen_phs_from_x_gpu call -- the momenta in the Whizard
 * layout [n_events][n_out][4] (the reference returns its per-branch array and gathers the
 * Whizard layout on the host), the beam energies E1 [n_events] followed by E2 [n_events],
 * and the ok flags [n_events].  NULL before the first call. */
double* get_momentum_device_pointer();
double* get_in_momentum_device_pointer();
bool* get_ok_device_pointer();
#endif

#endif /* GPU_PHS_H */
