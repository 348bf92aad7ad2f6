// Onesweep pass instantiations for u64 keys (every tile shape, direction,
// values flag and status-word width).
#include "radix_onesweep.hpp"

namespace clc {
namespace radix {
template void launch_onesweep<uint64_t, dev::KEY_UNSIGNED>(int, bool, bool, bool, const onesweep_args&,
                                           hipStream_t);
} // namespace radix
} // namespace clc
