// C ABI and host orchestration of one clustering run.
//
// The reference drives its path from main() (main.cu:35-459): device reset,
// ten cudaMallocs, one cudaMalloc per tree node, a host loop that copies the
// seed lists, collision matrix and all labels back and forth every
// iteration (main.cu:369-403, dbscan.cu:25-187).  Here a handle owns a
// device arena that grows on demand and is reused; a run is a fixed sequence
// of kernel launches on the handle's stream with two small host reads (the
// sort key width after the grid set-up, and the correction-queue length
// after each expansion pass).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "dclust.h"
#include "internal.h"
#include "kernels.h"

using namespace dclust;

namespace {

constexpr int64_t kChunk = 4096;  // points per chunk of bbox / cell-key blocks
constexpr int kMaxExpandPasses = 100000;
// heavy cells from which their adjacency pass runs beside the light one
constexpr long long kHeavyOverlap = 1024;

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

enum Ev {
  E_START, E_ROWS0, E_INDEX, E_CORE0, E_CORE, E_ADJ, E_PASS0, E_EXPAND, E_MERGE, E_FAR, E_BORDER,
  E_END, E_N
};

}  // namespace

struct dclust_handle {
  int device = 0;
  int num_cu = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority, like the first) for kernels that run
  // beside the main sequence, and the events that fork and join it
  hipStream_t stream2 = nullptr;
  hipEvent_t fork[4] = {};
  hipEvent_t ev[E_N] = {};
  std::vector<Buf> bufs;
  int64_t last_n = -1;
  uint8_t* last_core = nullptr;        // device, cell-sorted core flags of the last run
  const int32_t* last_orig = nullptr;  // device, cell-sorted input indices
  uint8_t* last_core_out = nullptr;    // device, input-order staging
  int last_err = 0;
  // launch shapes found by dclust_autotune_device, keyed by (dimension mix,
  // log2 of the total point count, precision form)
  std::map<std::tuple<int, int, int>, dclust_params> tuned;
  // replay of repeated identical calls
  std::vector<uint8_t> prev_sig;   // launch state of the last eager run
  int prev_npass = 0;              // its expansion passes
  long long prev_nlines = 0;       // its neighbour-line count
  long long prev_nheavy = 0;       // heavy cells of the last run (any kind)
  int prev_sort_passes = 0;        // its radix passes
  hipGraphExec_t gexec = nullptr;  // captured run
  std::vector<uint8_t> gsig;       // launch state it was captured for
  int gpasses = 0;                 // expansion passes it holds
  bool gphases = false;            // it records every phase event
  bool greadback = false;          // it ends with the counters' copies to `pin`
  // pinned host copy of the control counters ([C_NCTL]) and their stripes
  // ([C_NCTL][kStripes]), allocated with the handle (null: pageable copies)
  long long* pin = nullptr;
  long long* pin_dev = nullptr;  // its device address
  bool cap_phases = false;         // (set while a capture runs)
  bool graph_broken = false;       // capture failed on this device: eager only
  bool graph_off = false;          // eager only (while autotuning)
  uint8_t* scratch = nullptr;      // this device's scratch array (no-allocation runs)
  // the instance table and chunk list last copied to the device, and where
  // (a replay with the same bytes at the same place skips the copies)
  std::vector<uint8_t> up_bytes;
  const void* up_inst = nullptr;
  const void* up_chunks = nullptr;

  template <class T>
  T* get(int slot, size_t count, hipError_t* err) {
    if ((size_t)slot >= bufs.size()) bufs.resize(slot + 1);
    Buf& b = bufs[slot];
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (b.cap < bytes) {
      if (b.p) (void)hipFree(b.p);
      b.p = nullptr;
      b.cap = 0;
      size_t want = bytes + bytes / 8;  // headroom for slightly larger runs
      hipError_t e = hipMalloc(&b.p, want);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMalloc(&b.p, bytes);
        want = bytes;
      }
      if (e != hipSuccess) {
        *err = e;
        b.p = nullptr;
        return nullptr;
      }
      b.cap = want;
    }
    return (T*)b.p;
  }
};

namespace {

enum Slot {
  S_XYZ, S_INST, S_CHUNK, S_BBOX, S_CTL, S_KEY0, S_KEY1, S_VAL0, S_VAL1, S_HIST, S_OFFS,
  S_DTOT, S_PASS, S_CSTART, S_CKEY, S_CINST, S_CM, S_ROWS, S_CELLOF, S_X0, S_X1, S_X2,
  S_G0, S_G1, S_G2, S_ORIG, S_CORE, S_LABEL, S_PARENT, S_COMPMIN, S_PAIRS, S_Q0, S_Q1,
  S_KEYV, S_FLAG, S_RANK, S_SCAN, S_LABOUT, S_COREOUT, S_PAIRCNT, S_SEGPOS, S_LINES, S_AMASK,
  S_HEAVY, S_REC, S_RECN, S_CXYZ, S_CCOMP, S_ADJPTS, S_BORDERPTS, S_CORECELLS, S_TUNELAB, S_STAT, S_GLINES,
  S_CZ, S_XOFF, S_OS_LABEL, S_OS_STATE, S_OS_SEEDS, S_OS_LEN, S_OS_CONTACT, S_OS_FIRST, S_ROWSFB
};

// line table entries (8 bytes each) above which neighbour lines are found
// by binary search instead
constexpr long long kMaxLineTable = 1LL << 27;  // 1 GiB at most
// expansion passes a captured run may hold; runs that need more stay eager
constexpr int kGraphMaxPasses = 8;

// Device memory of dclust_run_device_noalloc: one array per GPU, placed with
// the code object when it is loaded, so a run needs no allocation call.  In
// host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (scratch_address) and only
// that address is handed to kernels.  Every no-allocation run on a GPU uses
// the same array, so such runs on one GPU must not overlap.  The kernels
// write every word they read before reading it (the run's own init kernels
// clear counters, lists and forests), so nothing is cleared here.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxScratchDevices] = {};

// Device address of this device's scratch array (the device must be
// current); looked up on the first call per device.
hipError_t scratch_address(int device, uint8_t** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = (uint8_t*)p;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

// Where a run's device buffers come from: the handle's arena (grown with
// hipMalloc and reused), or, for the no-allocation entry point, a fixed
// region handed out by a bump pointer in 256-byte steps.  A region too small
// for the run sets `over` and hands out nothing.
struct RunMem {
  dclust_handle* h;
  uint8_t* base = nullptr;  // region; null = the arena
  size_t cap = 0;
  size_t used = 0;
  bool over = false;
  hipError_t err = hipSuccess;
  template <class T>
  T* get(int slot, size_t count) {
    if (!base) return h->get<T>(slot, count, &err);
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    const size_t off = (used + 255) & ~(size_t)255;
    if (off > cap || bytes > cap - off) {
      over = true;
      return nullptr;
    }
    used = off + bytes;
    return (T*)(base + off);
  }
};

int check_params(const dclust_params* p) {
  if (!p) return DCLUST_E_PARAM;
  if (p->dim != 2 && p->dim != 3) return DCLUST_E_PARAM;
  if (!(p->eps > 0.0) || !std::isfinite(p->eps)) return DCLUST_E_PARAM;
  if (p->minpts < 1) return DCLUST_E_PARAM;
  if (p->partitions_r < 0 || p->seed_list_size < 0 || p->blocks < 0 || p->block_threads < 0 ||
      p->extra_collision < 0)
    return DCLUST_E_PARAM;
  // THREAD_COUNT: any workgroup size up to 1024 is accepted (common.h:14
  // uses 16); the expansion runs it rounded up to whole wavefronts, at most
  // 256 threads (effective_threads)
  if (p->block_threads > 1024) return DCLUST_E_PARAM;
  // seed lists live in LDS: (threads / 64) * seeds * 4 bytes <= 64 KiB
  if (p->seed_list_size > 4096) return DCLUST_E_PARAM;
  if (p->precision != DCLUST_PRECISION_FP64 && p->precision != DCLUST_PRECISION_FP32_GUARD)
    return DCLUST_E_PARAM;
  return DCLUST_OK;
}

float ev_ms(dclust_handle* h, int a, int b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, h->ev[a], h->ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.f;
  }
  return ms;
}

struct RunShape {
  int seeds;
  int threads;
  int blocks;
  int waves;
  int pair_per;
  int budget;
  int64_t pair_cap;
};

// Workgroup size the expansion runs for a THREAD_COUNT of t (0 = auto):
// rounded up to whole wavefronts of 64 lanes (each wavefront owns a seed
// list, so a partial wavefront would idle lanes the hardware runs anyway),
// at most 256 (the kernel's launch bound).
int effective_threads(int t) {
  if (t <= 0) return 256;
  return std::min(256, (t + 63) / 64 * 64);
}

// Built-in launch shapes of the expansion kernel, by the dimension mix of
// the call and its total point count: the fastest of dclust_autotune_device's
// 24 candidate shapes on one MI355X (256 CUs) for representative problems of
// each class (`bench.py --tune-table`; the measurements are recorded with
// the design notes).  One expansion launch serves every instance of a batch,
// so a batch is one problem of its total size.
struct ShapeRow {
  int dims;       // 1: 2-D only, 2: 3-D only, 3: both
  int log2_max;   // rows apply up to this floor(log2(points))
  int seeds, threads, per_cu;
};
constexpr ShapeRow kShapeTable[] = {
    {1, 16, 256, 128, 4},  {1, 19, 2048, 128, 4}, {1, 62, 1024, 256, 4},
    {2, 16, 256, 128, 4},  {2, 19, 2048, 128, 4}, {2, 22, 1024, 256, 4}, {2, 62, 1024, 128, 4},
    {3, 22, 1024, 256, 8}, {3, 62, 512, 256, 8},
};

const ShapeRow* shape_row(int dims, int64_t n) {
  int lg = 0;
  while (lg < 62 && (1LL << (lg + 1)) <= n) ++lg;
  for (const ShapeRow& r : kShapeTable)
    if (r.dims == dims && lg <= r.log2_max) return &r;
  return nullptr;
}

// Launch shape: explicit parameters win; a call that leaves seed list,
// workgroup size and workgroup count all at 0 takes the built-in table's
// shape for its dimension mix and size, other zeros the rule (seed list 1024
// entries, 256 threads, blocks at resident occupancy, at most 8 per CU).
RunShape pick_shape(dclust_handle* h, const dclust_params& p, int64_t n, bool f32, int dims) {
  RunShape r;
  const ShapeRow* row = (p.seed_list_size == 0 && p.block_threads == 0 && p.blocks == 0)
                            ? shape_row(dims, n)
                            : nullptr;
  r.seeds = p.seed_list_size > 0 ? p.seed_list_size : (row ? row->seeds : 1024);
  r.threads = row ? row->threads : effective_threads(p.block_threads);
  if (p.blocks > 0) {
    r.blocks = p.blocks;
  } else if (row) {
    // as measured: workgroups beyond those resident start as others finish
    r.blocks = h->num_cu * row->per_cu;
  } else {
    int per_cu = expand_blocks_per_cu(r.seeds, r.threads);
    r.blocks = h->num_cu * std::min(per_cu, 8);
  }
  r.blocks = std::max(8, (r.blocks + 7) / 8 * 8);  // multiple of 8 (XCD remap)
  // EXTRA_COLLISION_SIZE: collision pairs buffered per workgroup; split over
  // its wavefronts (each wavefront is a chain owner)
  const int waves_per_block = r.threads / 64;
  r.waves = r.blocks * waves_per_block;
  r.pair_per = p.extra_collision > 0 ? std::max(1, p.extra_collision / waves_per_block) : 512;
  r.pair_cap = (int64_t)r.pair_per * r.waves;
  // cells a wave expands per pass before handing its list on (a long chain
  // is shared out by the next pass; concurrent seeding already splits large
  // clusters between waves, so the budget rarely binds)
  r.budget = 4096;
  if (const char* e = std::getenv("DCLUST_EXPAND_BUDGET")) r.budget = std::max(1, std::atoi(e));
  (void)n;
  (void)f32;
  return r;
}

bool graph_enabled(const dclust_handle* h) {
  if (h->graph_off || h->graph_broken) return false;
  const char* e = std::getenv("DCLUST_GRAPH");
  return !(e && e[0] == '0');
}

void drop_graph(dclust_handle* h) {
  if (h->gexec) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipGraphExecDestroy(h->gexec);
  }
  h->gexec = nullptr;
  h->gsig.clear();
  h->gpasses = 0;
}

// Phase times of a replay: each event recorded inside a graph is a node that
// leaves the device idle for about 5.5 us.  The run's start and end events
// are recorded on the stream around the graph launch, outside the graph; a
// captured run holds only the two events of the neighbour-count kernel,
// unless the environment variable DCLUST_PHASE_TIMES=1 asks for every phase.
bool phase_times_requested() {
  const char* e = std::getenv("DCLUST_PHASE_TIMES");
  return e && e[0] == '1';
}
bool outside_graph(int e) { return e == E_START || e == E_END; }
bool always_in_graph(int e) { return e == E_CORE0 || e == E_CORE; }

// Everything a launch sequence depends on besides the coordinates: the
// device state (sizes, buffer and input/output pointers, switches), the
// launch shape, the precision form and the instance and chunk tables.
std::vector<uint8_t> launch_signature(const Dev& d, const RunShape& r, bool f32,
                                      const std::vector<Inst>& inst,
                                      const std::vector<Chunk>& chunks) {
  std::vector<uint8_t> v;
  auto put = [&v](const void* p, size_t nb) {
    const uint8_t* b = (const uint8_t*)p;
    v.insert(v.end(), b, b + nb);
  };
  put(&d, sizeof(d));
  const int64_t sh[9] = {r.seeds,    r.threads, r.blocks,  r.waves,
                         r.pair_per, r.budget,  r.pair_cap, f32 ? 1 : 0,
                         phase_times_requested() ? 1 : 0};
  put(sh, sizeof(sh));
  put(inst.data(), sizeof(Inst) * inst.size());
  put(chunks.data(), sizeof(Chunk) * chunks.size());
  return v;
}

// Phase event; inside a stream capture it becomes a node of the graph.
hipError_t mark(dclust_handle* h, int e, bool graph, hipStream_t s = nullptr) {
  if (graph && (outside_graph(e) || (!h->cap_phases && !always_in_graph(e)))) return hipSuccess;
  if (!s) s = h->stream;
  return graph ? hipEventRecordWithFlags(h->ev[e], s, hipEventRecordExternal)
               : hipEventRecord(h->ev[e], s);
}

// Work on the second stream after everything enqueued so far on the first
// (fork event k), and the first stream's wait for it (join event k + 1).
hipError_t fork_second(dclust_handle* h, int k) {
  hipError_t e = hipEventRecord(h->fork[k], h->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(h->stream2, h->fork[k], 0);
  return e;
}
hipError_t join_second(dclust_handle* h, int k) {
  hipError_t e = hipEventRecord(h->fork[k + 1], h->stream2);
  if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->fork[k + 1], 0);
  return e;
}

// Keys, sort, CSR, gather and line table, then the neighbour rows (one
// stream: the gather fills the device, so running the rows beside it gains
// nothing).
hipError_t enqueue_index(dclust_handle* h, const Dev& d, int64_t n, int nchunks, int grid_cells,
                         int passes, bool graph) {
  const hipStream_t s = h->stream;
  hipError_t e;
  launch_index_sort(d, n, nchunks, passes, h->num_cu, s);
  launch_gather(d, n, s);
  launch_lines(d, grid_cells, s);
  if ((e = mark(h, E_ROWS0, graph)) != hipSuccess) return e;
  launch_rows(d, n, grid_cells, s);
  return hipSuccess;
}

// counters, bounding boxes, grid set-up
void enqueue_head(const Dev& d, int nchunks, hipStream_t s) {
  launch_init_head(d, s);
  launch_index_head(d, nchunks, s);
}

// per-cell state, core flags, near adjacency, seeding segments
hipError_t enqueue_core(dclust_handle* h, const Dev& d, const RunShape& shape, bool f32,
                        bool graph) {
  const hipStream_t s = h->stream;
  hipError_t e;
  if ((e = mark(h, E_INDEX, graph)) != hipSuccess) return e;
  launch_init_run(d, shape.waves, s);
  // the core (neighbour-count) kernel is bracketed by its own events
  if ((e = mark(h, E_CORE0, graph)) != hipSuccess) return e;
  launch_core(d, f32, s);
  if ((e = mark(h, E_CORE, graph)) != hipSuccess) return e;
  // heavy cells on the second stream beside the light ones when the last
  // run had many (the fork and join cost more than a near-empty heavy pass
  // takes); the labels do not depend on it
  if (h->prev_nheavy >= kHeavyOverlap) {
    if ((e = fork_second(h, 2)) != hipSuccess) return e;
    launch_adj_heavy(d, f32, h->num_cu * 16, h->stream2);
    launch_adj_light(d, f32, s);
    if ((e = join_second(h, 2)) != hipSuccess) return e;
  } else {
    launch_adj(d, f32, h->num_cu * 16, s);
  }
  if ((e = mark(h, E_ADJ, graph)) != hipSuccess) return e;
  return hipSuccess;
}

// `check` > 0: the last pass of a captured run, whose carry kernel also
// marks the run abandoned when work is still pending after it
void enqueue_pass(const Dev& d, const RunShape& shape, int pass, hipStream_t s, int check = 0) {
  // (pass 0 needs no preparation: the counters it would reset are zero from
  // the start of the run; it drains no queue, so nothing is carried)
  if (pass > 0) launch_pass_prep(d, pass, s);
  launch_expand(d, pass, shape.blocks, shape.seeds, shape.threads, shape.budget, s);
  if (pass > 0 || check > 0) launch_pass_carry(d, pass, check, s);
}

// chain merge, far pairs, border points, canonical labels
hipError_t enqueue_tail(dclust_handle* h, const Dev& d, bool f32, int grid_cells, bool graph) {
  const hipStream_t s = h->stream;
  hipError_t e;
  if ((e = mark(h, E_EXPAND, graph)) != hipSuccess) return e;
  launch_merge(d, grid_cells, s);
  if ((e = mark(h, E_MERGE, graph)) != hipSuccess) return e;
  launch_far(d, f32, grid_cells, s);
  if ((e = mark(h, E_FAR, graph)) != hipSuccess) return e;
  launch_border(d, f32, s);
  if ((e = mark(h, E_BORDER, graph)) != hipSuccess) return e;
  launch_relabel(d, d.n, s);
  if ((e = mark(h, E_END, graph)) != hipSuccess) return e;
  return hipGetLastError();
}

// The control counters and their stripes into the handle's pinned buffer
// (inside a capture this becomes the graph's last kernel).
hipError_t enqueue_readback(dclust_handle* h, const Dev& d, hipStream_t s) {
  launch_readback(d, h->pin_dev, s);
  return hipGetLastError();
}

// A whole run with no host read in between, captured into a graph:
// `sort_passes` radix passes (the key width of the previous run; a wider key
// abandons the run on the device) and `passes` expansion passes (a pass with
// nothing pending exits at once).
hipError_t capture_run(dclust_handle* h, const Dev& d, const RunShape& shape, bool f32,
                       int nchunks, int grid_cells, int sort_passes, int passes,
                       hipGraphExec_t* out) {
  const hipStream_t s = h->stream;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  h->cap_phases = phase_times_requested();
  // (the start and end events are recorded on the stream around each
  // launch, outside the graph: the start event is processed while the host
  // is still submitting the graph)
  enqueue_head(d, nchunks, s);
  hipError_t body = enqueue_index(h, d, d.n, nchunks, grid_cells, sort_passes, true);
  if (body == hipSuccess) body = enqueue_core(h, d, shape, f32, true);
  // work still pending after the last pass: the rest of the graph returns at
  // once and the host runs the call eagerly
  for (int p = 0; body == hipSuccess && p < passes; ++p) {
    enqueue_pass(d, shape, p, s, p == passes - 1 ? passes : 0);
    if (p == 0) body = mark(h, E_PASS0, true);
  }
  if (body == hipSuccess) body = enqueue_tail(h, d, f32, grid_cells, true);
  if (body == hipSuccess && h->pin) body = enqueue_readback(h, d, s);
  e = hipStreamEndCapture(s, &g);
  if (body != hipSuccess) e = body;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  return e;
}

// Control counters and their striped parts, after the stream drains
// (`enqueued`: the copies to the pinned buffer are already on the stream).
hipError_t read_counters(dclust_handle* h, const Dev& d, hipStream_t s, long long* ctl_h,
                         std::vector<long long>& stat_h, bool enqueued = false) {
  hipError_t e = hipSuccess;
  if (h->pin) {
    if (!enqueued) e = enqueue_readback(h, d, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
      std::memcpy(ctl_h, h->pin, sizeof(long long) * C_NCTL);
      std::memcpy(stat_h.data(), h->pin + C_NCTL, sizeof(long long) * C_NCTL * kStripes);
    }
    return e;
  }
  e = hipMemcpyAsync(ctl_h, d.ctl, sizeof(long long) * C_NCTL, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(stat_h.data(), d.stat, sizeof(long long) * stat_h.size(),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

// Per-instance error codes of the grid set-up; *first = the first non-zero.
hipError_t read_inst_errors(const Dev& d, int32_t ninst, int32_t* inst_err, int* first) {
  std::vector<Inst> got(ninst);
  hipError_t e = hipMemcpy(got.data(), d.inst, sizeof(Inst) * ninst, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  *first = 0;
  for (int32_t i = 0; i < ninst; ++i) {
    if (inst_err) inst_err[i] = got[i].err;
    if (got[i].err && !*first) *first = got[i].err;
  }
  return hipSuccess;
}

// The launch shape a run used, as the common.h names count it.
void set_shape_stats(const RunShape& shape, dclust_stats* stats) {
  stats->eff_blocks = shape.blocks;
  stats->eff_block_threads = shape.threads;
  stats->eff_seed_list_size = shape.seeds;
  stats->eff_extra_collision = shape.pair_per * (shape.threads / 64);
}

// Records the run's device state for dclust_last_core_flags and fills the
// statistics (counters summed over their stripes, phase times from events).
void finish_run(dclust_handle* h, const Dev& d, const RunShape& shape, long long* ctl_h,
                const std::vector<long long>& stat_h, int npass, bool replayed,
                dclust_stats* stats) {
  for (int c = 0; c < C_NCTL; ++c)
    for (int k = 0; k < kStripes; ++k) ctl_h[c] += stat_h[(size_t)c * kStripes + k];
  h->last_n = d.n;
  h->prev_nheavy = ctl_h[C_NHEAVY];
  h->last_core = d.core;
  h->last_orig = d.orig;
  h->last_core_out = d.core_out;
  if (!stats) return;
  // a replay of a graph without the phase events reports those phases as 0
  const bool all_ev = !replayed || h->gphases;
  auto phase_ms = [&](int a, int b) { return all_ev ? (double)ev_ms(h, a, b) : 0.0; };
  stats->t_index_ms = ev_ms(h, E_START, E_CORE0);
  stats->t_rows_ms = phase_ms(E_ROWS0, E_INDEX);
  stats->n_core_cells = ctl_h[C_NCORECELLS];
  stats->n_adj_points = ctl_h[C_NADJPTS];
  stats->n_lines = ctl_h[C_NLINES];
  stats->work[0] = ctl_h[C_CORE_GROUPS];
  stats->work[1] = ctl_h[C_CORE_LINES];
  stats->work[2] = ctl_h[C_ADJ_LINES];
  stats->work[3] = ctl_h[C_ADJ_CELLS];
  stats->t_core_ms = ev_ms(h, E_CORE0, E_CORE);
  stats->t_adj_ms = phase_ms(E_CORE, E_ADJ);
  stats->t_expand_ms = phase_ms(E_ADJ, E_EXPAND);
  stats->t_merge_ms = phase_ms(E_EXPAND, E_MERGE);
  stats->t_far_ms = phase_ms(E_MERGE, E_FAR);
  stats->t_border_ms = phase_ms(E_FAR, E_BORDER);
  stats->t_relabel_ms = phase_ms(E_BORDER, E_END);
  stats->t_total_ms = ev_ms(h, E_START, E_END);
  stats->key_bits = (int32_t)ctl_h[C_KEYBITS];
  stats->n_cells = ctl_h[C_NCELL];
  stats->n_core = ctl_h[C_NCORE];
  stats->n_noise = ctl_h[C_NNOISE];
  stats->n_clusters = ctl_h[C_NCLUST];
  stats->n_chains = ctl_h[C_NCHAIN];
  stats->n_collisions = ctl_h[C_NPAIR];
  stats->n_overflow = ctl_h[C_OVF];
  stats->n_passes = npass;
  stats->cand_total = ctl_h[C_CAND];
  stats->nbr_cells_total = ctl_h[C_NBRC];
  stats->band_rechecks = ctl_h[C_BAND];
  stats->band_flips = ctl_h[C_FLIP];
  stats->asym_pairs = ctl_h[C_ASYM];
  stats->pairs_evaluated = ctl_h[C_EVAL];
  stats->band_rechecks_core = ctl_h[C_BAND_CORE];
  stats->rec_overflow = ctl_h[C_RECOVF];
  stats->diag[0] = ctl_h[C_POPS];
  stats->diag[1] = ctl_h[C_DENSE];
  stats->diag[2] = ctl_h[C_NHEAVY];
  stats->diag[3] = ctl_h[C_EDGES];
  stats->diag[4] = (int64_t)(1000.0 * phase_ms(E_ADJ, E_PASS0));
  stats->diag[5] = ctl_h[C_BUSY_SUM];
  stats->diag[6] = ctl_h[C_BUSY_MAX];
  stats->diag[7] = shape.waves;
  stats->diag[8] = ctl_h[C_ADJ_TESTS];
  stats->diag[9] = ctl_h[C_ADJ_FREE];
  stats->diag[10] = ctl_h[C_FAR_TESTS];
  stats->diag[11] = ctl_h[C_FAR_JOINS];
  stats->diag[12] = ctl_h[C_POPS_MAX];
  stats->diag[13] = ctl_h[C_SEEDS_MAX];
  stats->diag[14] = replayed ? 1 : 0;
  stats->diag[15] = ctl_h[C_ROUNDS_MAX];
  set_shape_stats(shape, stats);
}

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) {                \
      herr = e_;                           \
      goto fail;                           \
    }                                      \
  } while (0)

// One run over `ninst` instances whose coordinates are already on the
// device (d_xyz, rows back to back) and whose labels go to d_labels.
int onestep_loop(dclust_handle* h, const Dev& d, const dclust_params& prm, int threads,
                 int32_t* d_labels, dclust_stats* stats);

// onestep_threads > 0: the one-step baseline with workgroups of that size
// runs after the index instead of the clustering phases
// region != null: every device buffer of the run comes from that region
// (cap bytes), nothing is allocated and the run is eager (no graph capture).
int run_device(dclust_handle* h, int32_t ninst, const double* d_xyz, const int64_t* n,
               const dclust_params* params, int32_t* d_labels, int32_t* inst_err,
               dclust_stats* stats, int onestep_threads = 0, uint8_t* region = nullptr,
               size_t region_cap = 0) {
  const bool onestep = onestep_threads > 0;
  const bool fixed_mem = region != nullptr;
  hipError_t herr = hipSuccess;
  int rc = DCLUST_OK;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (ninst <= 0) return DCLUST_E_PARAM;
  bool f32 = params[0].precision == DCLUST_PRECISION_FP32_GUARD;
  std::vector<Inst> inst(ninst);
  std::vector<Chunk> chunks;
  int64_t ntot = 0, off = 0;
  for (int32_t i = 0; i < ninst; ++i) {
    int c = check_params(&params[i]);
    if (c) return c;
    if (n[i] < 0) return DCLUST_E_PARAM;
    if ((params[i].precision == DCLUST_PRECISION_FP32_GUARD) != f32) return DCLUST_E_PARAM;
    Inst& I = inst[i];
    std::memset(&I, 0, sizeof(I));
    I.dim = params[i].dim;
    I.minpts = params[i].minpts;
    I.eps = params[i].eps;
    I.eps2 = params[i].eps * params[i].eps;
    I.r = params[i].partitions_r;
    I.pt_begin = ntot;
    I.pt_end = ntot + n[i];
    I.xyz_off = off;
    for (int64_t b = 0; b < n[i]; b += kChunk) {
      Chunk ch;
      ch.inst = i;
      ch.pad = 0;
      ch.begin = ntot + b;
      ch.end = ntot + std::min(n[i], b + kChunk);
      chunks.push_back(ch);
    }
    ntot += n[i];
    off += n[i] * I.dim;
  }
  // cell indices are packed into 28 bits in the neighbour-line table
  if (ntot >= kMaxCells) return DCLUST_E_CAPACITY;
  if (inst_err)
    for (int32_t i = 0; i < ninst; ++i) inst_err[i] = 0;
  if (stats) stats->n_points = ntot;
  if (ntot == 0) return DCLUST_OK;

  int dims = 0;
  for (int32_t i = 0; i < ninst; ++i) dims |= params[i].dim == 3 ? 2 : 1;
  RunShape shape = pick_shape(h, params[0], ntot, f32, dims);
  if (fixed_mem && params[0].extra_collision == 0) {
    // collision pairs beyond the per-wave buffer are merged inline: a
    // smaller buffer leaves more of the fixed region to the points
    shape.pair_per = 64;
    shape.pair_cap = (int64_t)shape.pair_per * shape.waves;
  }
  if (onestep) {  // the expansion buffers stay unused: keep them minimal
    shape.blocks = 8;
    shape.waves = 8 * (shape.threads / 64);
    shape.pair_cap = (int64_t)shape.pair_per * shape.waves;
  }
  const int nchunks = (int)chunks.size();
  const int grid_cells = h->num_cu * 8;
  const hipStream_t s = h->stream;

  Dev d;
  std::memset(&d, 0, sizeof(d));
  d.n = ntot;
  d.ninst = ninst;
  d.nchunks = nchunks;
  d.xyz = d_xyz;
  d.pair_cap = shape.pair_cap;
  d.pair_per = shape.pair_per;
  d.rstride = ntot;
  // neighbourhood sizes (stats.cand_total / nbr_cells_total) cost an extra
  // dependent load per stencil line; they are summed only on request
  {
    const char* e = std::getenv("DCLUST_NBR_STATS");
    d.nbr_stats = (e && e[0] == '1') ? 1 : 0;
  }
  // neighbour records of non-core points: a non-core point has fewer than
  // minpts - 1 neighbours outside its own cell (plus pairs whose box test is
  // asymmetric); up to kRecSlots are kept, a longer record is rescanned
  int rec_stride = 0;
  for (int32_t i = 0; i < ninst; ++i) rec_stride = std::max(rec_stride, params[i].minpts);
  rec_stride = std::min(rec_stride, kRecSlots);
  d.rec_stride = rec_stride;
  RunMem mem{h, region, region_cap};
  {
    RunMem& M = mem;
    d.inst = M.get<Inst>(S_INST, ninst);
    d.chunks = M.get<Chunk>(S_CHUNK, nchunks);
    d.bbox = M.get<unsigned long long>(S_BBOX, (size_t)ninst * kBoxStripes * 8);
    d.ctl = M.get<long long>(S_CTL, C_NCTL);
    d.stat = M.get<long long>(S_STAT, (size_t)C_NCTL * kStripes);
    d.key[0] = M.get<uint64_t>(S_KEY0, ntot);
    d.key[1] = M.get<uint64_t>(S_KEY1, ntot);
    d.val[0] = M.get<int32_t>(S_VAL0, ntot);
    d.val[1] = M.get<int32_t>(S_VAL1, ntot);
    d.hist = M.get<uint32_t>(S_HIST, sort_hist_entries(ntot));
    d.offs = M.get<uint32_t>(S_OFFS, sort_hist_entries(ntot));
    d.dtot = M.get<uint32_t>(S_DTOT, 256);
    d.pass_scatter = M.get<int32_t>(S_PASS, kMaxPasses);
    d.cell_start = M.get<int32_t>(S_CSTART, ntot + 1);
    d.cell_key = M.get<uint64_t>(S_CKEY, ntot);
    d.cell_xyz = M.get<uint64_t>(S_CXYZ, ntot);
    d.cell_z = M.get<int32_t>(S_CZ, ntot);
    d.cell_comp = M.get<int32_t>(S_CCOMP, ntot);
    // striped work lists: sub-list capacity = the most its workgroups append
    const int64_t cap_p = (point_blocks(ntot) + kStripes - 1) / kStripes * 256;
    const int64_t cap_c = (cell_comp_blocks(ntot) + kStripes - 1) / kStripes * 1024;
    d.adj_pts = SList{M.get<int32_t>(S_ADJPTS, (size_t)cap_p * kStripes),
                      (unsigned long long*)(d.stat + (int64_t)C_NADJPTS * kStripes), cap_p};
    d.border_pts = SList{M.get<int32_t>(S_BORDERPTS, (size_t)cap_p * kStripes),
                         (unsigned long long*)(d.stat + (int64_t)C_NBORDERPTS * kStripes), cap_p};
    d.heavy = SList{M.get<int32_t>(S_HEAVY, (size_t)cap_p * kStripes),
                    (unsigned long long*)(d.stat + (int64_t)C_NHEAVY * kStripes), cap_p};
    d.core_cells = SList{M.get<int32_t>(S_CORECELLS, (size_t)cap_c * kStripes),
                         (unsigned long long*)(d.stat + (int64_t)C_NCORECELLS * kStripes), cap_c};
    const int64_t cap_r = row_list_cap(ntot);
    d.rows_fb = SList{M.get<int32_t>(S_ROWSFB, (size_t)cap_r * kStripes),
                      (unsigned long long*)(d.stat + (int64_t)C_NROWSFB * kStripes), cap_r};
    d.cell_inst = M.get<int32_t>(S_CINST, ntot);
    d.cell_ncore = M.get<int32_t>(S_CM, ntot);
    d.rows = M.get<int32_t>(S_ROWS, (size_t)ntot * kMaxRows);
    d.cell_of = M.get<int32_t>(S_CELLOF, ntot);
    if (!f32) {  // the FP32 form reads FP64 coordinates from the caller's rows
      d.X[0] = M.get<double>(S_X0, ntot);
      d.X[1] = M.get<double>(S_X1, ntot);
      d.X[2] = M.get<double>(S_X2, ntot);
    } else {
      d.xoff = M.get<int64_t>(S_XOFF, ntot);
    }
    d.G[0] = M.get<uint32_t>(S_G0, ntot);
    d.G[1] = M.get<uint32_t>(S_G1, ntot);
    d.G[2] = M.get<uint32_t>(S_G2, ntot);
    d.orig = M.get<int32_t>(S_ORIG, ntot);
    d.core = M.get<uint8_t>(S_CORE, ntot);
    d.cell_label = M.get<int32_t>(S_LABEL, ntot);
    d.amask = M.get<unsigned long long>(S_AMASK, (size_t)ntot);
    d.rec_n = M.get<uint8_t>(S_RECN, ntot);
    if (rec_stride > 0) d.rec = M.get<int32_t>(S_REC, (size_t)ntot * rec_stride);
    d.parent = M.get<int32_t>(S_PARENT, ntot);
    d.compmin = M.get<int32_t>(S_COMPMIN, ntot);
    d.pairs = M.get<int2>(S_PAIRS, shape.pair_cap);
    d.pair_cnt = M.get<int32_t>(S_PAIRCNT, shape.waves);
    d.seg_pos = M.get<int64_t>(S_SEGPOS, shape.waves);
    d.q[0] = M.get<int32_t>(S_Q0, ntot);
    d.q[1] = M.get<int32_t>(S_Q1, ntot);
    d.keyv = M.get<int32_t>(S_KEYV, ntot);
    d.flag = M.get<uint8_t>(S_FLAG, ntot);
    d.rank = M.get<int32_t>(S_RANK, ntot + 1);
    d.scan_tmp = M.get<int64_t>(S_SCAN, scan_tmp_entries(ntot));
    d.core_out = M.get<uint8_t>(S_COREOUT, ntot);
    if (mem.over) return DCLUST_E_CAPACITY;  // the fixed region is too small
    if (mem.err != hipSuccess) {
      herr = mem.err;
      return DCLUST_E_OOM;
    }
  }
  d.labels_out = d_labels;

  // A call whose launch state equals the previous call's (sizes, parameters,
  // pointers) is captured once into a graph and replayed from then on.
  std::vector<uint8_t> sig;
  const bool graph_on = !onestep && !fixed_mem && graph_enabled(h);
  if (graph_on) sig = launch_signature(d, shape, f32, inst, chunks);
  // (a run in a fixed region leaves the handle's graph alone: destroying it
  // would free memory)
  if (!fixed_mem && h->gexec && (!graph_on || sig != h->gsig)) drop_graph(h);
  if (graph_on && !h->gexec && !h->prev_sig.empty() && sig == h->prev_sig &&
      h->prev_npass < kGraphMaxPasses) {
    Dev dg = d;
    // line table sized from the previous run; a larger line count falls back
    // to the binary search on the device
    const long long cap = std::min<long long>(
        kMaxLineTable, std::max<long long>(h->prev_nlines + h->prev_nlines / 4, 1024));
    hipError_t ae = hipSuccess;
    dg.lines = h->get<int2>(S_GLINES, (size_t)cap, &ae);
    if (dg.lines) {
      dg.line_cap = cap;
      const int gp = std::max(2, h->prev_npass + 1);
      hipGraphExec_t ge = nullptr;
      const int sp = std::max(1, std::min(kMaxPasses, h->prev_sort_passes));
      if (capture_run(h, dg, shape, f32, nchunks, grid_cells, sp, gp, &ge) == hipSuccess) {
        h->gexec = ge;
        h->gsig = sig;
        h->gpasses = gp;
        h->gphases = h->cap_phases;
        h->greadback = h->pin != nullptr;
      } else {
        (void)hipGetLastError();
        h->graph_broken = true;  // capture unsupported here: stay eager
      }
    }
  }

  {
    long long ctl_h[C_NCTL];
    std::vector<long long> stat_h((size_t)C_NCTL * kStripes);
    // the host-built instance table and chunk list (the kernels rewrite only
    // derived fields of the table, and k_init_head clears its error codes)
    std::vector<uint8_t> up(sizeof(Inst) * ninst + sizeof(Chunk) * nchunks);
    std::memcpy(up.data(), inst.data(), sizeof(Inst) * ninst);
    std::memcpy(up.data() + sizeof(Inst) * ninst, chunks.data(), sizeof(Chunk) * nchunks);
    auto upload = [&]() -> hipError_t {
      if (!fixed_mem && up == h->up_bytes && h->up_inst == d.inst && h->up_chunks == d.chunks)
        return hipSuccess;
      hipError_t ue =
          hipMemcpyAsync(d.inst, inst.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, s);
      if (ue == hipSuccess)
        ue = hipMemcpyAsync((void*)d.chunks, chunks.data(), sizeof(Chunk) * nchunks,
                            hipMemcpyHostToDevice, s);
      if (fixed_mem || ue != hipSuccess) {
        h->up_bytes.clear();  // (a fixed region is not the arena)
        h->up_inst = h->up_chunks = nullptr;
      } else {
        h->up_bytes = up;
        h->up_inst = d.inst;
        h->up_chunks = d.chunks;
      }
      return ue;
    };
    if (h->gexec && !fixed_mem) {
      HIPCHK(upload());
      HIPCHK(hipEventRecord(h->ev[E_START], s));
      HIPCHK(hipGraphLaunch(h->gexec, s));
      HIPCHK(hipEventRecord(h->ev[E_END], s));  // (after the counters' readback)
      HIPCHK(read_counters(h, d, s, ctl_h, stat_h, h->greadback));
      if (!ctl_h[C_ABANDON]) {
        if (ctl_h[C_ERR]) {
          int first = 0;
          HIPCHK(read_inst_errors(d, ninst, inst_err, &first));
          if (ninst == 1 || !inst_err) return first ? first : DCLUST_E_PARAM;
          rc = first;
        }
        finish_run(h, d, shape, ctl_h, stat_h, (int)std::max<long long>(1, ctl_h[C_LASTPASS]),
                   true, stats);
        return rc;
      }
      // the expansion needed more passes than the graph holds: run eagerly
      // (the next identical call captures a graph with more passes)
      drop_graph(h);
    }

    int npass = 0;
    HIPCHK(upload());
    HIPCHK(mark(h, E_START, false));
    enqueue_head(d, nchunks, s);
    HIPCHK(hipMemcpyAsync(ctl_h, d.ctl, sizeof(ctl_h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // more cells than the 63-bit sort key holds, over all instances
    if (ctl_h[C_KEYOVF]) {
      if (inst_err)
        for (int32_t i = 0; i < ninst; ++i) inst_err[i] = DCLUST_E_CAPACITY;
      return DCLUST_E_CAPACITY;
    }
    if (ctl_h[C_ERR]) {
      int first = 0;
      HIPCHK(read_inst_errors(d, ninst, inst_err, &first));
      if (ninst == 1 || !inst_err) return first ? first : DCLUST_E_PARAM;
      rc = first;  // batched with per-instance errors: keep going
    }
    const int passes = (int)((ctl_h[C_KEYBITS] + 7) / 8);
    const long long nlines = ctl_h[C_NLINES];
    if (stats) stats->key_bits = (int32_t)ctl_h[C_KEYBITS];
    // dense line table when it is small enough, else binary search
    d.line_cap = 0;
    if (nlines <= kMaxLineTable) {
      d.lines = mem.get<int2>(S_LINES, (size_t)std::max<long long>(nlines, 1));
      if (d.lines) {
        d.line_cap = nlines;
      } else if (!fixed_mem) {
        herr = mem.err;
        return DCLUST_E_OOM;
      }
      // (a fixed region without room for the table: the rows are found by
      // binary search over the cell keys instead)
    }
    HIPCHK(enqueue_index(h, d, ntot, nchunks, grid_cells, passes, false));
    if (onestep) {
      HIPCHK(mark(h, E_INDEX, false));
      return onestep_loop(h, d, params[0], onestep_threads, d_labels, stats);
    }
    HIPCHK(enqueue_core(h, d, shape, f32, false));

    // expansion passes until no seed is pending: pass p drains the queue of
    // pass p-1 (overflowed seed lists, lists left when a wave's pop budget
    // ran out) and continues every unfinished seeding segment
    for (npass = 0; npass < kMaxExpandPasses; ++npass) {
      enqueue_pass(d, shape, npass, s);
      if (npass == 0) HIPCHK(mark(h, E_PASS0, false));
      HIPCHK(hipMemcpyAsync(ctl_h, d.ctl, sizeof(ctl_h), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (ctl_h[C_Q0_N + ((npass + 1) & 1)] == 0 && ctl_h[C_SEGS_LEFT] == 0) break;
    }
    ++npass;
    HIPCHK(enqueue_tail(h, d, f32, grid_cells, false));
    HIPCHK(read_counters(h, d, s, ctl_h, stat_h));
    finish_run(h, d, shape, ctl_h, stat_h, npass, false, stats);
    if (npass >= kMaxExpandPasses) return DCLUST_E_CAPACITY;
    if (graph_on) {
      h->prev_sig.swap(sig);
      h->prev_npass = npass;
      h->prev_nlines = nlines;
      h->prev_sort_passes = passes;
    }
    return rc;
  }
fail:
  (void)hipGetLastError();
  if (stats) stats->hip_error = (int32_t)herr;
  h->last_err = (int)herr;
  h->prev_sig.clear();
  return DCLUST_E_HIP;
}

// One-step baseline (CUDA-DClust): after the index, one launch per step of
// every chain, the seed-list lengths read back after each launch, refills
// when every list is empty, one union of the recorded contacts at the end.
int onestep_loop(dclust_handle* h, const Dev& d, const dclust_params& prm, int threads,
                 int32_t* d_labels, dclust_stats* stats) {
#define OSCHK(x)                                 \
  do {                                           \
    const hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) {                      \
      (void)hipGetLastError();                   \
      if (stats) stats->hip_error = (int32_t)e_; \
      h->last_err = (int)e_;                     \
      return DCLUST_E_HIP;                       \
    }                                            \
  } while (0)
  const hipStream_t s = h->stream;
  const int64_t n = d.n;
  const int B = prm.blocks > 0 ? prm.blocks : 256;
  const int T = threads;
  const int S = prm.seed_list_size > 0 ? prm.seed_list_size : 128;
  hipError_t herr = hipSuccess;
  OneStep o;
  std::memset(&o, 0, sizeof(o));
  o.blocks = B;
  o.S = S;
  // contacts are folded into the host union-find once half the list is
  // used; one launch adds far fewer than the other half
  o.cap = (int32_t)std::min<int64_t>(std::max<int64_t>(2 * n, 1 << 21), INT32_MAX / 2);
  // a smaller list (a test knob: folds happen every few launches)
  if (const char* e = std::getenv("DCLUST_ONESTEP_CONTACTS")) {
    const long v = std::atol(e);
    if (v >= 64 && v < o.cap) o.cap = (int32_t)v;
  }
  o.label = h->get<int32_t>(S_OS_LABEL, n, &herr);
  o.state = h->get<int32_t>(S_OS_STATE, n, &herr);
  o.seeds = h->get<int32_t>(S_OS_SEEDS, (size_t)B * S, &herr);
  o.len = h->get<int32_t>(S_OS_LEN, (size_t)B + 2, &herr);  // lengths, then the counters
  o.contacts = h->get<int2>(S_OS_CONTACT, o.cap, &herr);
  int32_t* d_first = h->get<int32_t>(S_OS_FIRST, B, &herr);
  if (herr != hipSuccess) return DCLUST_E_OOM;
  o.ctr = o.len + B;

  std::vector<int32_t> orig(n), inv(n), lab(n), len(B + 1), first(B), par;
  std::vector<int2> ct;
  int64_t launches = 0, rounds = 0, ncontacts = 0;
  auto find = [&](int32_t x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  // the contacts recorded so far (cnt of them) into the union-find; the
  // device counter restarts at 0
  auto fold = [&](int32_t cnt) -> int {
    if (cnt > o.cap) return DCLUST_E_CAPACITY;
    ct.resize(cnt);
    if (cnt > 0 &&
        hipMemcpy(ct.data(), o.contacts, sizeof(int2) * cnt, hipMemcpyDeviceToHost) != hipSuccess)
      return DCLUST_E_HIP;
    const int32_t zero = 0;
    if (hipMemcpy(o.ctr, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
      return DCLUST_E_HIP;
    for (const int2& e : ct) {
      if (e.x < 0 || e.x >= (int64_t)par.size() || e.y < 0 || e.y >= (int64_t)par.size())
        return DCLUST_E_CAPACITY;
      const int32_t a = find(e.x), b = find(e.y);
      if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
    ncontacts += cnt;
    return DCLUST_OK;
  };
  long long ctl_h[C_NCTL];
  OSCHK(hipMemcpyAsync(ctl_h, d.ctl, sizeof(ctl_h), hipMemcpyDeviceToHost, s));
  OSCHK(hipMemcpyAsync(orig.data(), d.orig, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  launch_os_init(o, n, s);
  OSCHK(hipStreamSynchronize(s));
  if (ctl_h[C_ERR] || ctl_h[C_ABANDON]) return DCLUST_E_PARAM;
  for (int64_t i = 0; i < n; ++i) inv[orig[i]] = (int32_t)i;
  {
    // refill seeds: unprocessed points from the back of the input order
    int64_t cursor = n - 1;
    const int64_t max_launches = 4 * n + 1024;
    for (;;) {
      OSCHK(hipMemcpyAsync(lab.data(), o.label, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
      OSCHK(hipStreamSynchronize(s));
      int got = 0;
      for (int b = 0; b < B; ++b) {
        while (cursor >= 0 && lab[inv[cursor]] != -1) --cursor;
        first[b] = cursor >= 0 ? inv[cursor--] : -1;
        got += first[b] >= 0;
      }
      if (!got) break;
      if ((rounds + 1) * (int64_t)B >= INT32_MAX) return DCLUST_E_CAPACITY;
      OSCHK(hipMemcpyAsync(d_first, first.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, s));
      launch_os_refill(o, d_first, s);
      const int32_t chain0 = (int32_t)(rounds * B);
      ++rounds;
      const size_t old = par.size();
      par.resize((size_t)rounds * B);
      for (size_t c = old; c < par.size(); ++c) par[c] = (int32_t)c;
      for (;;) {
        launch_os_step(d, o, chain0, T, s);
        ++launches;
        OSCHK(hipMemcpyAsync(len.data(), o.len, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost,
                             s));
        OSCHK(hipStreamSynchronize(s));
        if (len[B] >= o.cap / 2) {
          const int rc = fold(len[B]);
          if (rc) return rc;
        }
        bool any = false;
        for (int b = 0; b < B && !any; ++b) any = len[b] > 0;
        if (!any) break;
        if (launches > max_launches) return DCLUST_E_CAPACITY;
      }
    }
  }
  launch_os_border(d, o, s);
  std::vector<int32_t> state(n);
  int32_t nct = 0;
  OSCHK(hipMemcpyAsync(lab.data(), o.label, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  OSCHK(hipMemcpyAsync(state.data(), o.state, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  OSCHK(hipMemcpyAsync(&nct, o.ctr, sizeof(nct), hipMemcpyDeviceToHost, s));
  OSCHK(hipStreamSynchronize(s));
  {
    const int rc = fold(nct);
    if (rc) return rc;
  }

  // canonical ids: a cluster is numbered by the rank of its smallest core
  // input index
  const int64_t nchains = rounds * B;
  std::vector<int32_t> cmin(nchains, INT32_MAX);
  int64_t ncore = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (state[i] == 0 || lab[i] == -1) return DCLUST_E_CAPACITY;  // never expanded
    if (state[i] == 2) {
      ++ncore;
      const int32_t r = find(lab[i]);
      cmin[r] = std::min(cmin[r], orig[i]);
    }
  }
  std::vector<std::pair<int32_t, int32_t>> mins;
  for (int64_t c = 0; c < nchains; ++c)
    if (par[c] == c && cmin[c] != INT32_MAX) mins.push_back({cmin[c], (int32_t)c});
  std::sort(mins.begin(), mins.end());
  std::vector<int32_t> id(nchains, 0);
  for (size_t k = 0; k < mins.size(); ++k) id[mins[k].second] = (int32_t)k + 1;
  std::vector<int32_t> out(n);
  int64_t nnoise = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t v = lab[i] >= 0 ? id[find(lab[i])] : 0;
    out[orig[i]] = v;
    nnoise += v == 0;
  }
  OSCHK(hipMemcpyAsync(d_labels, out.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  OSCHK(mark(h, E_END, false));
  OSCHK(hipStreamSynchronize(s));
  if (stats) {
    stats->n_points = n;
    stats->n_core = ncore;
    stats->n_noise = nnoise;
    stats->n_clusters = (int64_t)mins.size();
    stats->n_chains = nchains;
    stats->n_collisions = ncontacts;
    stats->n_passes = launches;
    stats->key_bits = (int32_t)ctl_h[C_KEYBITS];
    stats->eff_blocks = B;
    stats->eff_block_threads = T;
    stats->eff_seed_list_size = S;
    stats->eff_extra_collision = 0;
    stats->t_index_ms = ev_ms(h, E_START, E_INDEX);
    stats->t_expand_ms = ev_ms(h, E_INDEX, E_END);
    stats->t_total_ms = ev_ms(h, E_START, E_END);
  }
  return DCLUST_OK;
}
#undef OSCHK

}  // namespace

extern "C" {

int dclust_params_default(dclust_params* p) {
  if (!p) return DCLUST_E_PARAM;
  std::memset(p, 0, sizeof(*p));
  p->dim = 2;        // DIMENSION
  p->eps = 1.5;      // EPS
  p->minpts = 3;     // MINPTS
  p->partitions_r = 3;  // PARTITION_SIZE
  p->seed_list_size = 128;    // MAX_SEEDS
  p->blocks = 8;              // THREAD_BLOCKS
  p->block_threads = 16;      // THREAD_COUNT (runs as one wavefront of 64)
  p->extra_collision = 256;   // EXTRA_COLLISION_SIZE
  p->precision = DCLUST_PRECISION_FP64;
  return DCLUST_OK;
}

int dclust_create(dclust_handle** out, int device) {
  if (!out) return DCLUST_E_PARAM;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return DCLUST_E_NODEVICE;
  }
  if (device < 0 || device >= ndev) return DCLUST_E_PARAM;
  dclust_handle* h = new (std::nothrow) dclust_handle();
  if (!h) return DCLUST_E_OOM;
  h->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete h;
    return DCLUST_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete h;
    return DCLUST_E_HIP;
  }
  h->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // this device's scratch array for dclust_run_device_noalloc (the device is
  // current); without it that entry point reports DCLUST_E_HIP
  if (scratch_address(device, &h->scratch) != hipSuccess) {
    (void)hipGetLastError();
    h->scratch = nullptr;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess) {
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DCLUST_E_HIP;
  }
  for (auto& e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      dclust_destroy(h);
      return DCLUST_E_HIP;
    }
  for (auto& e : h->fork)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      dclust_destroy(h);
      return DCLUST_E_HIP;
    }
  // coherent and mapped: the readback kernel's stores are visible to the
  // host once the stream has drained
  if (hipHostMalloc((void**)&h->pin, sizeof(long long) * C_NCTL * (1 + kStripes),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&h->pin_dev, h->pin, 0) != hipSuccess) {
    (void)hipGetLastError();
    if (h->pin) (void)hipHostFree(h->pin);
    h->pin = h->pin_dev = nullptr;  // counters are then copied to pageable memory
  }
  *out = h;
  return DCLUST_OK;
}

void dclust_destroy(dclust_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  drop_graph(h);
  for (auto& b : h->bufs)
    if (b.p) (void)hipFree(b.p);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : h->fork)
    if (e) (void)hipEventDestroy(e);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->pin) (void)hipHostFree(h->pin);
  delete h;
}

int dclust_run_device(dclust_handle* h, const double* d_xyz, int64_t n, const dclust_params* p,
                      int32_t* d_labels_out, dclust_stats* stats) {
  if (!h || !p || n < 0 || (n > 0 && (!d_xyz || !d_labels_out))) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  return run_device(h, 1, d_xyz, &n, p, d_labels_out, nullptr, stats);
}

int dclust_run_device_noalloc(dclust_handle* h, const double* d_xyz, int64_t n,
                              const dclust_params* p, int32_t* d_labels_out,
                              dclust_stats* stats) {
  if (!h || !p || n < 0 || (n > 0 && (!d_xyz || !d_labels_out))) return DCLUST_E_PARAM;
  if (!h->scratch) return DCLUST_E_HIP;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  return run_device(h, 1, d_xyz, &n, p, d_labels_out, nullptr, stats, 0, h->scratch,
                    kScratchBytes);
}

int dclust_run_onestep_device(dclust_handle* h, const double* d_xyz, int64_t n,
                              const dclust_params* p, int32_t* d_labels_out, dclust_stats* stats) {
  if (!h || !p || n < 0 || (n > 0 && (!d_xyz || !d_labels_out))) return DCLUST_E_PARAM;
  if (p->block_threads < 0 || p->block_threads > 1024) return DCLUST_E_PARAM;
  dclust_params q = *p;
  q.block_threads = 0;  // the main path's limit (<= 256) does not apply here
  q.precision = DCLUST_PRECISION_FP64;
  int c = check_params(&q);
  if (c) return c;
  if (p->blocks > (1 << 16)) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  h->prev_sig.clear();
  // whole wavefronts, as THREAD_COUNT is counted on the main path
  const int threads = p->block_threads > 0 ? (p->block_threads + 63) / 64 * 64 : 256;
  return run_device(h, 1, d_xyz, &n, &q, d_labels_out, nullptr, stats, threads);
}

int dclust_run(dclust_handle* h, const double* xyz, int64_t n, const dclust_params* p,
               int32_t* labels_out, dclust_stats* stats) {
  if (!h || !p || n < 0 || (n > 0 && (!xyz || !labels_out))) return DCLUST_E_PARAM;
  int c = check_params(p);
  if (c) return c;
  if (n == 0) {
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return DCLUST_OK;
  }
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  hipError_t herr = hipSuccess;
  double* dx = h->get<double>(S_XYZ, (size_t)n * p->dim, &herr);
  int32_t* dl = h->get<int32_t>(S_LABOUT, n, &herr);
  if (!dx || !dl) return DCLUST_E_OOM;
  if (hipMemcpyAsync(dx, xyz, sizeof(double) * n * p->dim, hipMemcpyHostToDevice, h->stream) !=
      hipSuccess)
    return DCLUST_E_HIP;
  int rc = run_device(h, 1, dx, &n, p, dl, nullptr, stats);
  if (rc != DCLUST_OK) return rc;
  if (hipMemcpyAsync(labels_out, dl, sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream) !=
          hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return DCLUST_E_HIP;
  return DCLUST_OK;
}

int dclust_run_batched_device(dclust_handle* h, int32_t count, const double* d_xyz,
                              const int64_t* n, const dclust_params* params, int32_t* d_labels,
                              int32_t* inst_err, dclust_stats* stats) {
  if (!h || count <= 0 || !n || !params || !d_xyz || !d_labels) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  return run_device(h, count, d_xyz, n, params, d_labels, inst_err, stats);
}

int dclust_run_batched(dclust_handle* h, int32_t count, const double* const* xyz, const int64_t* n,
                       const dclust_params* params, int32_t* const* labels, int32_t* inst_err,
                       dclust_stats* stats) {
  if (!h || count <= 0 || !xyz || !n || !params || !labels) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  int64_t tot = 0, el = 0;
  for (int32_t i = 0; i < count; ++i) {
    int c = check_params(&params[i]);
    if (c) return c;
    if (n[i] < 0 || (n[i] > 0 && (!xyz[i] || !labels[i]))) return DCLUST_E_PARAM;
    tot += n[i];
    el += n[i] * params[i].dim;
  }
  if (tot == 0) {
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (inst_err)
      for (int32_t i = 0; i < count; ++i) inst_err[i] = 0;
    return DCLUST_OK;
  }
  hipError_t herr = hipSuccess;
  double* dx = h->get<double>(S_XYZ, el, &herr);
  int32_t* dl = h->get<int32_t>(S_LABOUT, tot, &herr);
  if (!dx || !dl) return DCLUST_E_OOM;
  int64_t e = 0;
  for (int32_t i = 0; i < count; ++i) {
    if (n[i] > 0 && hipMemcpyAsync(dx + e, xyz[i], sizeof(double) * n[i] * params[i].dim,
                                   hipMemcpyHostToDevice, h->stream) != hipSuccess)
      return DCLUST_E_HIP;
    e += n[i] * params[i].dim;
  }
  int rc = run_device(h, count, dx, n, params, dl, inst_err, stats);
  if (rc != DCLUST_OK && !(inst_err && rc < 0 && rc != DCLUST_E_HIP && rc != DCLUST_E_OOM))
    return rc;
  int64_t o = 0;
  for (int32_t i = 0; i < count; ++i) {
    if (n[i] > 0 && hipMemcpyAsync(labels[i], dl + o, sizeof(int32_t) * n[i],
                                   hipMemcpyDeviceToHost, h->stream) != hipSuccess)
      return DCLUST_E_HIP;
    o += n[i];
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) return DCLUST_E_HIP;
  return rc;
}

int dclust_last_core_flags(dclust_handle* h, uint8_t* core_out, int64_t n) {
  if (!h || !core_out || n != h->last_n || !h->last_core) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  launch_core_out(h->last_core, h->last_orig, n, h->last_core_out, h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    (void)hipGetLastError();
    return DCLUST_E_HIP;
  }
  if (hipMemcpy(core_out, h->last_core_out, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
    return DCLUST_E_HIP;
  return DCLUST_OK;
}

// Tuning-table key of a problem: dimension mix (2, 3, or 5 for both), size
// class, precision form.
static std::tuple<int, int, int> tune_key(int32_t count, const int64_t* n,
                                          const dclust_params* params) {
  int dims = 0;
  int64_t tot = 0;
  for (int32_t i = 0; i < count; ++i) {
    dims |= params[i].dim == 3 ? 2 : 1;
    tot += n[i];
  }
  int lg = 0;
  while ((1LL << (lg + 1)) <= tot) ++lg;
  const int dm = dims == 3 ? 5 : (dims == 2 ? 3 : 2);
  return std::make_tuple(dm, lg, params[0].precision);
}

int dclust_autotune_device(dclust_handle* h, int32_t count, const double* d_xyz, const int64_t* n,
                           dclust_params* params) {
  if (!h || count <= 0 || !d_xyz || !n || !params) return DCLUST_E_PARAM;
  if (hipSetDevice(h->device) != hipSuccess) return DCLUST_E_HIP;
  int64_t tot = 0;
  for (int32_t i = 0; i < count; ++i) {
    int c = check_params(&params[i]);
    if (c) return c;
    if (n[i] < 0) return DCLUST_E_PARAM;
    tot += n[i];
  }
  hipError_t herr = hipSuccess;
  int32_t* lab = h->get<int32_t>(S_TUNELAB, std::max<int64_t>(tot, 1), &herr);
  if (!lab) return DCLUST_E_OOM;
  // candidates: seed-list length x workgroup size x workgroups per CU; each
  // is run twice and scored by its faster run
  const int seeds_c[] = {256, 512, 1024, 2048};
  const int threads_c[] = {128, 256};
  const int per_cu_c[] = {4, 8, 16};
  std::vector<dclust_params> trial(params, params + count);
  // candidates are timed by eager runs (no capture per candidate shape)
  struct Eager {
    dclust_handle* h;
    ~Eager() { h->graph_off = false; }
  } eager{h};
  h->graph_off = true;
  double best = 1e30;
  dclust_params best_p = params[0];
  for (int sd : seeds_c)
    for (int th : threads_c)
      for (int pc : per_cu_c) {
        const int waves_per_cu = pc * th / 64;
        if (waves_per_cu > 32) continue;  // beyond what a CU holds
        for (auto& q : trial) {
          q.seed_list_size = sd;
          q.block_threads = th;
          q.blocks = h->num_cu * pc;
          q.extra_collision = 0;
        }
        double t = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
          dclust_stats st;
          int rc = run_device(h, count, d_xyz, n, trial.data(), lab, nullptr, &st);
          if (rc != DCLUST_OK) return rc;
          t = std::min(t, st.t_total_ms);
        }
        if (t < best) {
          best = t;
          best_p = trial[0];
        }
      }
  for (int32_t i = 0; i < count; ++i) {
    params[i].seed_list_size = best_p.seed_list_size;
    params[i].block_threads = best_p.block_threads;
    params[i].blocks = best_p.blocks;
    params[i].extra_collision = best_p.extra_collision;
  }
  h->tuned[tune_key(count, n, params)] = best_p;
  return DCLUST_OK;
}

int dclust_autotune(dclust_handle* h, int64_t n, dclust_params* p) {
  if (!h || !p) return DCLUST_E_PARAM;
  bool f32 = p->precision == DCLUST_PRECISION_FP32_GUARD;
  auto it = h->tuned.find(tune_key(1, &n, p));
  if (it != h->tuned.end()) {
    if (p->seed_list_size == 0) p->seed_list_size = it->second.seed_list_size;
    if (p->block_threads == 0) p->block_threads = it->second.block_threads;
    if (p->blocks == 0) p->blocks = it->second.blocks;
    return DCLUST_OK;
  }
  RunShape r = pick_shape(h, *p, n, f32, p->dim == 3 ? 2 : 1);
  if (p->seed_list_size == 0) p->seed_list_size = r.seeds;
  if (p->block_threads == 0) p->block_threads = r.threads;
  if (p->blocks == 0) p->blocks = r.blocks;
  if (p->extra_collision == 0) p->extra_collision = r.pair_per * (r.threads / 64);
  return DCLUST_OK;
}

const char* dclust_strerror(int code) {
  switch (code) {
    case DCLUST_OK: return "ok";
    case DCLUST_E_PARAM: return "invalid parameter or non-finite input";
    case DCLUST_E_BINWIDTH: return "cell width (min extent / r) is less than eps";
    case DCLUST_E_IO: return "dataset or label file error";
    case DCLUST_E_HIP: return "HIP runtime error";
    case DCLUST_E_OOM: return "out of memory";
    case DCLUST_E_CAPACITY: return "size limit exceeded";
    case DCLUST_E_NODEVICE: return "no usable GPU";
    default: return "unknown error";
  }
}

const char* dclust_version(void) { return "dclust-mi355x 0.1 (gfx950)"; }

}  // extern "C"
