// fcm_column32.hip -- the FP32 kernels of the column path (fcm_column.hpp),
// built in a translation unit of their own so that both precisions compile
// in parallel.
#include "fcm_column.hpp"

namespace fcm {

template hipError_t spread_zfwd<float>(const GridGeom&, const BlobData&, const double*, const FieldPtrs<float>&,
                                       const TwiddleSet<float>&, hipStream_t, bool);
template hipError_t zinv_interp<float>(const GridGeom&, const BlobData&, const FieldPtrs<float>&,
                                       const TwiddleSet<float>&, double*, const KTerm&, hipStream_t, hipEvent_t, int);

}  // namespace fcm
