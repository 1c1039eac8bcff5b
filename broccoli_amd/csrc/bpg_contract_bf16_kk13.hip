// Instantiations of the split-BF16 contraction for 13..16 k-steps of 16
// subjects (N up to 256).
#include "bpg_contract_bf16.h"

namespace bpg {
template hipError_t dispatch_bf16_pp<13>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<14>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<15>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<16>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
}  // namespace bpg
