// hostutil.h -- host-side helpers of the C ABI shared by its translation
// units (defined in capi.hip and batch_host.hip).
#pragma once

#include "tucu.h"
#include "sepfir.h"
#include "batch.h"

#include <hip/hip_runtime.h>
#include <string>
#include <vector>

namespace tucu {
namespace host {

int    fail(int code, const std::string& msg);
int    device_fail(hipError_t e, const char* what);
size_t dtype_size(int dt);
int    gauss_class(int r);
int    check_image(const tucu_image* im, const char* name, bool need_data);
int    gauss_args(const float* const lobes[3], const int n[3], const tucu_image* in,
                  const tucu_image* outs, unsigned mask, int shift, SepArgs& a, int& rmax, int& odt);
// Aliasing of a fused Gaussian call: TUCU_ERR_INVALID if two live outputs
// overlap, -1 if an output overlaps the input, TUCU_OK otherwise.
int    gauss_aliasing(const tucu_image* in, const tucu_image* outs, unsigned live, const int* r, bool shift);
bool   gauss_use_mfma();
// Kernel family of the batched kernels: BATCH_MFMA (matrix cores, the
// default) or BATCH_VALU (TUCU_GAUSS_PATH=valu), read at the call.  A batch or
// server records the family when it is built and keeps it.
enum { BATCH_MFMA = 0, BATCH_VALU = 1 };
int    batch_family();
// Batched kernel of family fam and tile-shape index thi (0 .. NBTH-1) for a
// type pair.
const BatchEntry& batch_entry(int fam, int thi, int in_dt, int out_dt);
int    make_tiles(std::vector<SepArgs>& jobs, const BatchEntry& e, std::vector<BatchTile>& tiles);
int    batch_grid(const BatchEntry& e, int ntiles);

}   // namespace host
}   // namespace tucu
