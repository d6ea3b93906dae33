// Allocation-free entry points of libtrident_hip.
//
// Some callers cannot allocate device memory while they run: tight serving
// loops, harnesses that forbid hipMalloc.  tr_spadd_into_noalloc is
// tr_spadd_into (the double-buffer add of sp_mma_hybrid,
// include/cusparse_helpers.cuh:708-739) with every temporary of the add
// (task table, per-row chunk counts, chunk rows and offsets, scan partials,
// validation bitmaps) carved from one file-scope device array instead of the
// context's pool.  Inputs and outputs are still the caller's.
//
// The array is kNoallocBytes per GPU, held for the life of the process by
// every GPU the library runs on.  Host code never uses the array's name as
// an address (on the host it is only a stand-in for the symbol): its device
// address is looked up once per device with hipGetSymbolAddress while that
// device is current (noalloc_init, called from tr_ctx_create), and kernels
// receive that address.  Every call on one GPU shares the array and it keeps
// the last call's contents, so everything that must start at zero is cleared
// by a fill kernel on the call's stream, and calls that overlap on one GPU
// (two streams or threads at once) race.
#include <mutex>
#include <string>

#include "tr_internal.h"

namespace {

constexpr size_t kNoallocBytes = TR_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char noalloc_scratch[kNoallocBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
std::string g_err[kMaxDevices];

}  // namespace

namespace tr {

void noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        g_err[device] = std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(e);
        return;
    }
    g_ptr[device] = p;
}

static char *noalloc_base(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    TR_REQUIRE(device >= 0 && device < kMaxDevices, "noalloc: device id out of range");
    if (!g_ptr[device])
        fail(TR_ERR_HIP, "noalloc: scratch of device " + std::to_string(device) + " unavailable: " +
                             (g_err[device].empty() ? std::string("not initialised") : g_err[device]));
    return (char *)g_ptr[device];
}

int64_t noalloc_spadd_into(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *C, tr_csx_t *out, int64_t capacity) {
    char *base = noalloc_base(ctx->device);
    // [status words | validation bitmap | add temporaries], 256-byte pieces
    const size_t st_bytes = 256;
    const size_t val_bytes = (validate_tmp_bytes(std::max(A->nnz, C->nnz)) + 255) & ~(size_t)255;
    const size_t add_bytes = spadd_fixed_bytes(A->nrows, A->nnz, C->nnz);
    const size_t need = st_bytes + val_bytes + add_bytes;
    if (need > kNoallocBytes)
        fail(TR_ERR_INVALID, "tr_spadd_into_noalloc: operands need " + std::to_string(need) +
                                 " bytes of scratch, more than the fixed " + std::to_string(kNoallocBytes) +
                                 " (use tr_spadd_into)");
    int32_t *st = (int32_t *)base;
    uint32_t *starts = (uint32_t *)(base + st_bytes);
    launch_fill_i32(ctx, st, 0, 2);
    launch_validate(ctx, A, st, starts);
    launch_validate(ctx, C, st + 1, starts);  // same stream: A's checks finish first
    const int64_t hs = read_scalar<int64_t>(ctx, (const int64_t *)st);
    const int32_t s0 = (int32_t)(hs & 0xffffffff), s1 = (int32_t)(hs >> 32);
    if ((s0 | s1) & 3) fail(TR_ERR_INVALID, "tr_spadd_into_noalloc: malformed CSR operand");
    if ((s0 | s1) & 4) fail(TR_ERR_INVALID, "tr_spadd_into_noalloc: operands need sorted unique columns");
    return spadd_into_fixed(ctx, A, C, out->ptr_vec, out->idx_vec, out->val, capacity, base + st_bytes + val_bytes,
                            add_bytes);
}

}  // namespace tr
