// Prefix sums for CSR row pointers (hipCUB device scan).
#include <hipcub/hipcub.hpp>

#include "common.h"
#include "internal.h"

namespace hm {

int exclusive_offsets(const int64_t* in, int64_t* out, long long n, int64_t* total,
                      hipStream_t s) {
    HM_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    size_t bytes = 0;
    HM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out + 1, (int)n, s));
    void* tmp = nullptr;
    HM_TRY(dev_alloc(&tmp, bytes));
    hipError_t e = hipcub::DeviceScan::InclusiveSum(tmp, bytes, in, out + 1, (int)n, s);
    if (e == hipSuccess) e = hipMemcpyAsync(total, out + n, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(tmp);
    if (e != hipSuccess) {
        set_error("prefix sum failed: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}

}  // namespace hm
