// Slice passes at the transform lengths 5 x 2^k (80, 160, 320, 640, 1280):
// the kernels of slice_passes.hpp with launch variant 0, instantiated in a
// unit of their own so that they compile in parallel with kernels.hip.
#include "slice_passes.hpp"

namespace mtg {

template <class Z>
hipError_t launch_row_pass_m5(int n, const RowArgsT<Z> &a, hipStream_t s)
{
    switch (n) {
    case 80: return row_launch<80>(a, s);
    case 160: return row_launch<160>(a, s);
    case 320: return row_launch<320>(a, s);
    case 640: return row_launch<640>(a, s);
    case 1280: return row_launch<1280>(a, s);
    }
    return hipErrorInvalidValue;
}

template <class Z>
hipError_t launch_col_pass_m5(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s)
{
    if (probe && det) return hipErrorInvalidValue;
    switch (n) {
    case 80: return col_launch<80>(a, probe, det, s);
    case 160: return col_launch<160>(a, probe, det, s);
    case 320: return col_launch<320>(a, probe, det, s);
    case 640: return col_launch<640>(a, probe, det, s);
    case 1280: return col_launch<1280>(a, probe, det, s);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_row_pass_m5<float2>(int, const RowArgsT<float2> &, hipStream_t);
template hipError_t launch_row_pass_m5<double2>(int, const RowArgsT<double2> &, hipStream_t);
template hipError_t launch_col_pass_m5<float2>(int, const ColArgsT<float2> &, bool, bool, hipStream_t);
template hipError_t launch_col_pass_m5<double2>(int, const ColArgsT<double2> &, bool, bool, hipStream_t);

} // namespace mtg
