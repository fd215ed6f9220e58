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
bool   gauss_use_mfma();
// Batched kernel of tile-shape index thi (0 .. NBTH-1) for a type pair, from
// the kernel family in use (matrix-core by default).
const BatchEntry& batch_entry(int thi, int in_dt, int out_dt);
int    make_tiles(std::vector<SepArgs>& jobs, const BatchEntry& e, std::vector<BatchTile>& tiles);
int    batch_grid(const BatchEntry& e, int ntiles);

}   // namespace host
}   // namespace tucu
