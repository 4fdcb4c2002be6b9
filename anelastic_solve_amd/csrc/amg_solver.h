// amg_solver.h -- host orchestration of the multigrid pressure solve on one
// GPU: device state per instance, V-cycle step lists over a set of
// instances (one launch per level step), the FP64 outer loop.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "amg.h"
#include "amg_hierarchy.h"
#include "amg_internal.h"

namespace amg {

void set_error(const std::string& msg);
const std::string& last_error();

constexpr int kMaxInstances = 1024;

#define AMG_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
      return AMG_EHIP;                                                       \
    }                                                                        \
  } while (0)

// One work-list entry: descriptor index and its block count.
struct Item {
  int idx;
  int blocks;
};

// Owns device allocations.
struct DevArena {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  void* alloc(size_t n);
  void release();
  ~DevArena() { release(); }
};

// One Poisson problem resident on the device.
struct Instance {
  int id = 0;
  int gslot = 0;                 // slot of its gate in Context::d_igate
  Hierarchy H;
  int nx = 0, ny = 0, nz = 0;
  bool fp64 = false;             // V-cycle precision (true: plain FP64)
  std::vector<int> lev;          // DevLevel index per level, finest first
  std::vector<int> xf;           // DevXfer index per transition
  int l0 = 0;                    // first level of the coarse tail
  int ld = -1;                   // level solved directly (separable solve), or -1
  size_t dlds = 0;               // dynamic LDS of its plane transforms
  // the first level not cycled with one launch per step
  int lc() const { return ld >= 0 ? ld : l0; }
  int tail = -1;                 // DevTail index
  size_t tail_lds = 0;           // dynamic LDS bytes of the tail
  int tail_nzc = 16;             // tail kernel variant: column length class
  bool tail_222 = true;          //   and uniform 2:1 transfers
  bool tail_split = false;       // the coarse solve is its own kernel (large coarsest level)
  int fine[2] = {-1, -1};        // FP64 finest entries (u = p_cur for cur = 0/1)
  bool pairok = false;           // may pair cycles (DevLevel::pairok)
  int cur = 0;                   // which fine entry holds the current pressure
  std::vector<int> pre, post;
  DevArena mem;
  // projection metrics (ghost-inclusive): rho, rdxc, rdyc, rdzc, rdxe, rdye, rdze
  double* d_met = nullptr;
  // per-solve state
  bool converged = false, done = false;
  int iters = 0;
  int last_iters = 0;            // cycles of the previous solve (the gated batch size)
  double bnorm = 0.0, relres = 0.0, relmax = 0.0, res0 = 0.0;
  int status = AMG_OK;
  bool gate_seen = false;        // the host has read the gate of the first defect
  int units = 0;                 // corrections replayed from the gate
  int dev_cycles = 0;            // V-cycles the device had run at the last look
  size_t bytes_per_cycle = 0;    // a V-cycle and its FP64 correction
  size_t bytes_vcycle = 0, bytes_corr = 0;   // the two parts
  double* d_phi = nullptr;       // caller arrays of the current solve
  const double* d_rhs = nullptr;
  size_t device_bytes() const { return mem.bytes; }
};

// A set of instances sharing one stream and one descriptor table.
struct Context {
  amg_config cfg{};
  int device = 0;                // the HIP device current at init (the stream's)
  hipStream_t stream = nullptr;
  std::vector<Instance*> inst;
  std::vector<DevLevel> hL;
  std::vector<DevXfer> hX;
  std::vector<DevTail> hT;
  DevLevel* dL = nullptr;
  DevXfer* dX = nullptr;
  DevTail* dT = nullptr;
  size_t capL = 0, capX = 0, capT = 0;
  double* d_partial = nullptr;   // block partials
  size_t cap_partial = 0;
  double* d_out = nullptr;       // per-item results
  double* h_out = nullptr;       // pinned mirror
  ItemGate* d_igate = nullptr;   // per-instance gates, kMaxInstances
  ItemGate* h_igate = nullptr;   // and their pinned mirror
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // gated single-instance graphs of kChainLen[q] cycles
  static constexpr int kChainLen[3] = {1, 2, 4};
  hipGraphExec_t gchain[3] = {};
  void drop_graphs();
  bool graph_ok = true;

  int init(const amg_config& c);
  ~Context();
  // builds the hierarchy (or takes `built`, from check_grid) and allocates
  int add_instance(const amg_grid& g, Instance** out, Hierarchy* built = nullptr);
  int upload_tables();
  int pull_state();
  int solve(const std::vector<Instance*>& S, amg_stats* st);
  // pieces
  void vcycle(const std::vector<Instance*>& S, bool fp64);
  void zline_launches(const std::vector<Item>& items, bool fp64, int parity, int zg, int form = 0);
  void rb_launches(const std::vector<Item>& items, bool fp64, int parity, int zg);
  void direct_launches(const std::vector<Item>& planes, const std::vector<Item>& modes, size_t lds,
                       bool fp64);
  // the start of a solve, enqueued only: the gates of S are initialised on
  // the device, the host reads them with the first cycles' (sync_gates)
  int prepare_async(const std::vector<Instance*>& S);
  int gate_slots() const;
  // copies the gates to the host, waits, and replays those of S
  int sync_gates(const std::vector<Instance*>& S, double* bytes);
  const double* null_sums(const std::vector<Instance*>& S);
  int finish(const std::vector<Instance*>& S);
  int cycle_and_norms(const std::vector<Instance*>& S);
  // replays the corrections that instance I made on the device since the
  // host's last look (its gate as copied after a batch or round), adding
  // their bytes; AMG_EHIP when the device and the host disagree
  int absorb_gate(Instance* I, const ItemGate& g, double* bytes);
  // *tail: the solve's copy out and state reset were enqueued behind its
  // last batch (nothing is left for the caller to do but read the stats)
  int gated_cycles(Instance* I, double* bytes, bool* tail);
  void reset_gates(const std::vector<Instance*>& S, bool only_stopped);
  void post_norms(const std::vector<Instance*>& S, int m);
  int run_kernel(Instance* I, int which, int level, int reps, double* bytes);
  int level_field(Instance* I, int level, int field, int dir, double* buf);
  void remove_instance(Instance* I);
};

// Host-side validation of a grid under cfg (the checks of
// mgMK/mg_check_size.F90:43-151 and this build's own limits), building its
// hierarchy; no device work.  AMG_OK or an error code with the message set.
int check_grid(const amg_config& cfg, const amg_grid& g, Hierarchy& H);

// Algorithmic bytes of one V-cycle and its FP64 correction; X holds the
// instance's transfers, X[l] from level l to level l+1.
size_t cycle_bytes(const Instance& I, const std::vector<DevXfer>& X);
// Blocks of the instance's largest finest-level launch that writes block
// partials (each writes at most 4 values per block).
size_t partial_blocks(const Instance& I);

}  // namespace amg
