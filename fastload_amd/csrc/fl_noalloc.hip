// fl_spmv_assign_noalloc: y = A x like fl_spmv_assign, for callers that cannot allocate device
// memory during the call (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). No reference counterpart.
//
// fl_spmv_assign needs device memory beyond its inputs and outputs in three places: the PB
// schedule's product array (pb_ring products, allocated by the library on the matrix's first PB
// call), the HYB schedule's heavy-part y (m values) and its sub-matrix's product array, and the
// FAST kernel's split-BF16 work buffer (2*m floats, passed by the caller). Here all of them live
// in one file-scope device array, `scratch`, of kScratchBytes, which every GPU the library runs
// on holds until the process ends. Its device address is looked up once per device (the
// first call on that device) with hipGetSymbolAddress and kept; host code never hands the array's
// own name to a kernel (on the host that name is only a stand-in, not a device address).
//
// Every call on a GPU shares that GPU's array and the array keeps the last call's contents: the
// PB products are all written by phase 1 before phase 2 reads them, the HYB heavy kernel writes
// every row's value before the sub-matrix's phase 2 reads it, and the split-BF16 partial sums are
// cleared by a fill kernel, so nothing depends on what an earlier call left. Calls that
// overlap on one GPU (several streams, several host threads) race on the array.
#include <mutex>
#include <vector>

#include "fl_matrix.h"

static constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
std::vector<void *> g_scratch_ptr;  // per device: the device address of `scratch`

// Device address of this device's scratch array (looked up on the device's first call).
void *scratch_address() {
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        FL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            fl::set_error("fl_spmv_assign_noalloc: no device address for the scratch array");
            throw fl::HipError{hipErrorInvalidSymbol};
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

using namespace fl;

extern "C" {

int fl_spmv_assign_noalloc(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *stream) {
    FL_API_BEGIN
    if (!A || (A->tilenum && !d_sortx) || (A->m && !d_y)) {
        set_error("fl_spmv_assign_noalloc: NULL");
        return FL_ERR_INVALID;
    }
    // bytes of scratch this call needs; refuse before anything is enqueued
    size_t need = 0;
    if (kernel == FL_KERNEL_PB && A->pb_ready)
        need = (size_t)A->pb_ring * (val_kind == FL_VAL_F64 ? sizeof(double) : sizeof(float));
    else if (kernel == FL_KERNEL_HYB && A->hyb_ready)
        need = hyb_scratch_bytes(A, val_kind);
    else if (kernel == FL_KERNEL_FAST && val_kind == FL_VAL_BF16X2)
        need = sizeof(float) * 2 * (size_t)A->m;
    if (need > kScratchBytes) {
        set_error("fl_spmv_assign_noalloc: needs %zu bytes of scratch, more than the %zu-byte (64 MiB) static array; "
                  "use fl_spmv_assign",
                  need, kScratchBytes);
        return FL_ERR_INVALID;
    }
    void *p = scratch_address();
    return spmv(A, kernel, val_kind, d_sortx, d_y, p, true, (hipStream_t)stream, p);
    FL_API_END
}

}  // extern "C"
