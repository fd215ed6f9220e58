// Matrix-core Gaussian kernels reading f32 images.
#define IN_T float
#define IN_IDX 1
#include "inst_mfma.inc"

#ifdef TUCU_TIMELINE
// Reads / clears the phase timeline of the one-round launch (timeline build only).
extern "C" __attribute__((visibility("default"))) int tucu_timeline_read(void* host, size_t bytes)
{
    if (bytes > sizeof(tucu::tucu_timeline)) return -1;
    return int(hipMemcpyFromSymbol(host, HIP_SYMBOL(tucu::tucu_timeline), bytes, 0, hipMemcpyDeviceToHost));
}
extern "C" __attribute__((visibility("default"))) int tucu_timeline_clear()
{
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(tucu::tucu_timeline)) != hipSuccess) return -1;
    return int(hipMemset(p, 0, sizeof(tucu::tucu_timeline)));
}
#endif
