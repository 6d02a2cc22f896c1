// Slice passes at the transform lengths 3 x 2^k (96, 192, 384, 768, 1536):
// the kernels of slice_passes.hpp with launch variant 0, instantiated in a
// unit of their own so that they compile in parallel with kernels.hip.
#include "slice_passes.hpp"

namespace mtg {

template <class Z>
hipError_t launch_row_pass_m3(int n, const RowArgsT<Z> &a, hipStream_t s)
{
    switch (n) {
    case 96: return row_launch<96>(a, s);
    case 192: return row_launch<192>(a, s);
    case 384: return row_launch<384>(a, s);
    case 768: return row_launch<768>(a, s);
    case 1536: return row_launch<1536>(a, s);
    }
    return hipErrorInvalidValue;
}

template <class Z>
hipError_t launch_col_pass_m3(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s)
{
    if (probe && det) return hipErrorInvalidValue;
    switch (n) {
    case 96: return col_launch<96>(a, probe, det, s);
    case 192: return col_launch<192>(a, probe, det, s);
    case 384: return col_launch<384>(a, probe, det, s);
    case 768: return col_launch<768>(a, probe, det, s);
    case 1536: return col_launch<1536>(a, probe, det, s);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_row_pass_m3<float2>(int, const RowArgsT<float2> &, hipStream_t);
template hipError_t launch_row_pass_m3<double2>(int, const RowArgsT<double2> &, hipStream_t);
template hipError_t launch_col_pass_m3<float2>(int, const ColArgsT<float2> &, bool, bool, hipStream_t);
template hipError_t launch_col_pass_m3<double2>(int, const ColArgsT<double2> &, bool, bool, hipStream_t);

} // namespace mtg
