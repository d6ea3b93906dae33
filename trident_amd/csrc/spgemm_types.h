// Types shared by the multiply pipeline (k_spgemm.hip) and the batched
// stage executor (k_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>


namespace tr {

// tuning constants
constexpr int kLightMaxFCap = 1024; // largest light/heavy cut (light tables reach 2048 slots)
constexpr int kNumLight = 6;       // light classes: T = 64 << c, c < kNumLight
constexpr int kHeavyClass = kNumLight;      // heavy rows with more than kMidMaxF multiplications
constexpr int kMidClass = kNumLight + 1;    // heavy rows up to kMidMaxF multiplications
constexpr int kNumClasses = kNumLight + 2;
constexpr int kMidMaxF = 16384;
constexpr int kMidPassBits = 19;            // mid rows: at most 2^19-column passes, two blocks per CU
constexpr int kMidThreads = 512;
constexpr int kPassBits = 20;      // heavy bitmap covers 2^20 columns per pass
constexpr int kFineShift = 7;      // fine block = 128 columns
constexpr int kWinLog = 19;                                // window width cap: 2^19 columns
constexpr int kWinMaxWidth = 1 << kWinLog;                 // (column offsets stay below 2^20)
constexpr int kWidthShift = kWinLog - kFineShift;          // fine blocks per width quantum
constexpr int kBmWidth = 32768;                            // columns ranked per presence-bitmap pass
constexpr int kBmWords = kBmWidth / 64;                    // presence words per pass
static_assert(kWinMaxWidth <= (1 << 20) && kPassBits <= 20, "window column offsets fit 20 bits");
constexpr int kSparseAvg = 192;    // a pass with <= this many entries per width quantum is sparse
constexpr int kSparseH = 256;      // sparse windows: count quantum (<= kSparseH + 127 entries, any width)
constexpr int kHeavyThreads = 1024;  // symbolic heavy block
constexpr int kLongRow = 2048;       // heavy rows with more A entries take shorter items
constexpr int kCompactMax = 128;     // one-pass heavy rows with <= this many entries take the light numeric
static_assert(kCompactMax * 2 <= (64 << 2), "compact rows fit the light tables of classes 0..2");

struct Window {
    int32_t c0, c1;   // column range [c0, c1)
    int32_t off;      // offset of the window's first entry within the row
    int32_t cnt;      // entries in the window
};

struct Item {
    int32_t row;
    int32_t wbeg, wend;  // window index range
    int32_t pad;
};

struct Panel {
    int64_t nnz_begin;   // global offset of the panel's first entry
    int32_t row_begin;
    int32_t pad;
    int32_t *idx;
    double *val;
};

// One product C_t = A_t * B_t of a batch (a plain call is a batch of one).
// Rows of all tasks form one global row space: task t owns global rows
// [row_base, row_base + m).
struct TaskDev {
    const int32_t *ap, *ai;
    const double *av;
    const int32_t *bp, *bi;
    const double *bv;
    int32_t m;         // rows of A_t
    int32_t n;         // columns of B_t
    int32_t row_base;  // first global row
    int32_t pad;
};

// device counters written by the symbolic phase
struct Counters {
    int32_t cls[kNumClasses];        // rows per class
    unsigned long long flops;        // total multiplications
    unsigned long long heavy_flops;  // multiplications of heavy rows
    int32_t windows;                 // windows appended
    int32_t items;                   // items appended
    int32_t overflow;                // capacity exceeded
    int32_t pad1;
};


// Buffers of one pipeline: the context's scratch slots for plain calls, or
// owned allocations for batches (a captured graph needs fixed addresses).
struct Workspace {
    tr_ctx *ctx = nullptr;
    bool owned = false;
    std::vector<void *> p;
    std::vector<size_t> sz;
    std::vector<size_t> req;  // largest request per slot (sizes the memory bound of a call)
    Workspace(tr_ctx *c, bool own)
        : ctx(c), owned(own), p(S_NUM_SLOTS, nullptr), sz(S_NUM_SLOTS, 0), req(S_NUM_SLOTS, 0) {}
    void *get(int slot, size_t bytes);
    void release();
    ~Workspace() { release(); }
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
};

// One multiply job over a table of tasks.
struct MulPlan {
    int nt = 0;
    int32_t rows_total = 0;
    std::vector<TaskDev> h_tasks;
    TaskDev *d_tasks = nullptr;
    int64_t *flops = nullptr, *rownnz = nullptr, *rowptr = nullptr;
    int32_t *lists = nullptr;
    Counters *cnt = nullptr;
    Counters h{};
    Window *wins = nullptr;
    Item *items = nullptr;
    int32_t win_cap = 0, item_cap = 0;
    int32_t *clen = nullptr;
    int64_t *cbase = nullptr;
    int4 *cursors = nullptr;  // (cursor, column at the cursor, end of B_k) per A entry of multi-window items
    int32_t *order = nullptr;          // heavy items sorted by the column of their first window
    unsigned long long *order_hist = nullptr;
    int64_t cursor_cap = 0;
    int64_t *scan_tmp = nullptr;  // owned scan temp (graph mode)
    size_t scan_tmp_bytes = 0;
    tr_tuning_t tune{};           // the context's tuning when the plan was set up
};

// One add X_t + Y_t of a batch.  X rows come from xp (int32, per task) or,
// when xp64 is set, from a 64-bit pointer over the global row space.
struct AddTaskDev {
    const int32_t *xp;
    const int64_t *xp64;
    const int32_t *xi;
    const double *xv;
    const int32_t *yp;
    const int32_t *yi;
    const double *yv;
    int32_t *optr;     // output row pointer (m + 1), rebased to the task
    int32_t m;
    int32_t row_base;
};

struct AddPlan {
    int nt = 0;
    int32_t rows_total = 0;
    std::vector<AddTaskDev> h;
    AddTaskDev *d = nullptr;
    int32_t *nch = nullptr, *crow = nullptr, *ccount = nullptr;
    int64_t *cstart = nullptr, *coff = nullptr;
    int64_t chunk_cap = 0, nchunks = 0;
    int64_t *scan_tmp = nullptr;
    // replay guard (batches): device word set on a mismatch with the plan,
    // and the planned first output entry of every task (nt + 1 entries)
    int32_t *guard = nullptr;
    const int64_t *expect = nullptr;
};

void add_setup(tr_ctx *ctx, Workspace &ws, AddPlan &ap, const std::vector<AddTaskDev> &tasks, int64_t nnz_x,
               int64_t nnz_y);
void add_count(tr_ctx *ctx, AddPlan &ap, bool capture);
int64_t add_total(tr_ctx *ctx, AddPlan &ap);
void add_write(tr_ctx *ctx, AddPlan &ap, int32_t *oi, double *ov);

// Pipeline steps.  In eager mode the host reads counts between steps to size
// grids and lists; in capture mode every size comes from a previous eager run
// on the same operands.
void mul_setup(tr_ctx *ctx, Workspace &ws, MulPlan &pl, const std::vector<TaskDev> &tasks);
void mul_analyze(tr_ctx *ctx, MulPlan &pl);
void mul_read_counts(tr_ctx *ctx, MulPlan &pl);
void mul_symbolic(tr_ctx *ctx, Workspace &ws, MulPlan &pl, bool emit, bool capture);
int64_t mul_cursor_prep(tr_ctx *ctx, Workspace &ws, MulPlan &pl, bool capture);  // returns cursor total (eager)
void mul_numeric(tr_ctx *ctx, MulPlan &pl, Panel *d_panels, int np);
void mul_panel_bases(tr_ctx *ctx, MulPlan &pl, Panel *d_panels, int np);
void mul_check(tr_ctx *ctx, MulPlan &pl);  // reads the consistency flags (eager)

}  // namespace tr
