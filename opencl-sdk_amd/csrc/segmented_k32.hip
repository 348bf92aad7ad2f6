// Segmented sort instantiations for 32-bit keys (uint, int, float).
#include "segmented_kernels.hpp"

namespace clc {
namespace seg {
template std::size_t sort_typed<uint32_t, dev::KEY_UNSIGNED>(context&, bool, void*, bool, void*,
                                                   std::size_t,
                                                      const uint64_t*, std::size_t, bool,
                                                      const table&, bool);
template std::size_t sort_typed<uint32_t, dev::KEY_SIGNED>(context&, bool, void*, bool, void*,
                                                   std::size_t,
                                                    const uint64_t*, std::size_t, bool,
                                                    const table&, bool);
template std::size_t sort_typed<uint32_t, dev::KEY_FLOAT>(context&, bool, void*, bool, void*,
                                                   std::size_t,
                                                   const uint64_t*, std::size_t, bool,
                                                   const table&, bool);
} // namespace seg
} // namespace clc
