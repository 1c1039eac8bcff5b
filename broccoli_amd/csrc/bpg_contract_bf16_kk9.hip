// Instantiations of the split-BF16 contraction for 9..12 k-steps of 16
// subjects (N up to 192).
#include "bpg_contract_bf16.h"

namespace bpg {
template hipError_t dispatch_bf16_pp<9>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<10>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<11>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<12>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
}  // namespace bpg
