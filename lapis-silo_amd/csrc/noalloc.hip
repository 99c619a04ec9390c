// Device scratch of the allocation-free Mutations entry point (rdb_mutations_query_noalloc).
//
// One file-scope device array, `scratch`, holds everything that entry point's kernels need beyond
// their inputs and outputs: the staged filter program, the filter's row words and cardinalities,
// the Mutations kernels' partial sums, heavy sums, position-block totals, sparse accumulators and
// the rows before they are copied back.  Each GPU the library runs on holds its own copy for the
// life of the process.  Its device address comes from hipGetSymbolAddress, resolved once per
// device with that device current (the array's name is only a host-side stand-in); all host code
// uses that address.  Every call on one GPU shares the array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace rdb {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

hipError_t noallocScratch(int device, void** ptr) {
   if (device < 0 || device >= kMaxDevices || ptr == nullptr) return hipErrorInvalidValue;
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (g_scratch_ptr[device] == nullptr) {
      int current = 0;
      hipError_t err = hipGetDevice(&current);
      if (err != hipSuccess) return err;
      if (current != device && (err = hipSetDevice(device)) != hipSuccess) return err;
      void* p = nullptr;
      err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (current != device) (void)hipSetDevice(current);
      if (err != hipSuccess) return err;
      if (p == nullptr) return hipErrorInvalidSymbol;
      g_scratch_ptr[device] = p;
   }
   *ptr = g_scratch_ptr[device];
   return hipSuccess;
}

}  // namespace rdb
