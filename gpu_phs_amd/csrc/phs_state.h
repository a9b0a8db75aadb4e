// Library-internal state shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpu_phs.h"
#include "phs_kernels.h"
#include "tree_compiler.h"

namespace gphs {

constexpr int64_t kMergeBelowDefault = (int64_t)1 << 18;

// ---- errors: every entry point records its status for gpu_phs_last_error()
int set_error(int code, const std::string& msg);
inline int ok_status() { return set_error(GPU_PHS_OK, ""); }
int hip_error(hipError_t e, const char* what);

std::mutex& api_mutex();

struct ProcessObj {
  CompiledProcess cp;
};

constexpr int kTileClasses = 2;
constexpr int kMaxStreams = 4;

// A run of consecutive tiles of one tile class whose channels share one tree shape (or -1:
// channels without a shape).
using ShapeRun = TileRun;

struct PlanObj {
  std::vector<gpu_phs_segment> user;       // segments as given
  std::vector<const double*> beam_e1, beam_e2;   // per-event beam energies of each segment, or null
  std::vector<double*> seg_msq;            // per-branch msq output of each segment, or null
  int specialize = 1;                      // 1: trees with a shape run the shape-specialised kernels
  // a shape whose tiles in a class hold fewer events than this runs in the class's one launch
  // of the generic kernel instead of a launch of its own (a launch of a few tiles costs about
  // one wave's lifetime however small it is)
  int64_t merge_below = kMergeBelowDefault;
  std::vector<ShapeRun> shape_runs[2];     // each tile class split by shape (empty: one generic launch)
  std::vector<ChannelDesc> chans;
  std::vector<NodeCmd> nodes;
  std::vector<SegDesc> segs;
  std::vector<TileDesc> tiles;
  std::vector<int> seg_tile_begin, seg_tile_end;  // tile ranges per segment (segment order)
  int n_banks = 0;
  // Tile classes: channels of up to 8 randoms whose only deferring branch is the root (class 0)
  // and the rest (class 1) run as separate launches, each with its own kernel variant and
  // launch shape.  Class c's tiles are [class_begin[c], class_end[c]) of the tile table.
  // defaults from measurement: the small programs with their randoms in LDS, 256 threads; the
  // others with 64 threads and 4 events per thread
  LaunchConfig lcs[kTileClasses] = {LaunchConfig{256, 1, 0, 0, 2}, LaunchConfig{64, 4, 0, 0, 0}};
  DevicePlan dps[kTileClasses];
  int class_begin[kTileClasses] = {0, 0}, class_end[kTileClasses] = {0, 0};
  std::vector<int> chan_class;   // class of each plan channel
  int order = 0;
  int serial_classes = 0;                   // 1: the two classes run one after the other on one stream
  void* dev = nullptr;                      // one allocation: tiles | segs | chans | nodes
  size_t dev_bytes = 0;
  DevicePlan dp;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t capture_stream = nullptr;
  // the second tile class runs beside the first: a side stream joined by events
  hipStream_t side_stream = nullptr;
  hipEvent_t fork_event = nullptr, join_event = nullptr;
  // n_streams > 1: every launch of the plan (a shape run, or a class's generic launch) goes to
  // one of n_streams streams -- the caller's and n_streams - 1 side streams joined back by
  // events -- heaviest first onto the least loaded one, so that a launch's last workgroups
  // share the GPU with other launches instead of leaving it partly idle
  int n_streams = 1;
  hipStream_t lane_streams[kMaxStreams - 1] = {};
  hipEvent_t lane_join[kMaxStreams - 1] = {};
  struct Job {
    int cls, begin, end, sid;
    double bytes;   // algorithmic bytes of its events
  };
  std::vector<Job> jobs;   // every launch of a run of the plan, heaviest first
  // adjoint: a copy of the tables with the gradient buffers, per-tile partials, tiles by channel
  void* adj_dev = nullptr;
  size_t adj_bytes = 0;
  std::vector<SegDesc> adj_segs;
  DevicePlan adj_dp;
  AdjLayout adj_lay{};
  int adj_part_stride = 0;
  double* adj_part = nullptr;
  int* adj_chan_tiles = nullptr;
  int* adj_chan_off = nullptr;
  double* adj_ranges = nullptr;
  unsigned int* adj_done = nullptr;   // per-channel count of finished reduction ranges (0 between launches)
  int adj_sub = 1;            // 64-event sub-tiles per tile (one wavefront each)
  bool adj_valid = false;     // the device copies above match adj_segs
  std::vector<int> chan_process, chan_channel;   // plan channel -> (process handle, channel)
  // inverse mapping: the process's whole channel table and the per-segment rows, on the device
  void* inv_dev = nullptr;
  size_t inv_bytes = 0;
  std::vector<InvSeg> inv_segs;
  int inv_process = -1;
  InverseTables inv;
  ~PlanObj();
};

// Registry (guarded by api_mutex()).
int register_process(std::unique_ptr<ProcessObj> p);
ProcessObj* find_process(int h);
int drop_process(int h);
int register_plan(std::unique_ptr<PlanObj> p);
PlanObj* find_plan(int h);
int drop_plan(int h);

// Build a plan's tables (host side) and upload them.  Returns a status.
int plan_build(PlanObj& pl);
int plan_run(PlanObj& pl, hipStream_t stream);
int plan_run_plain(PlanObj& pl, hipStream_t stream);
// Inverse mapping of the plan's events into every channel of their process (block 0 = 128).
int plan_inverse(PlanObj& pl, const int64_t* seg_row0, int64_t n_rows, double* factors, double* x_out, int block,
                 hipStream_t stream);

// Device check used by every entry point that touches the GPU.
int require_device();

}  // namespace gphs
