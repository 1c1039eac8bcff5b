// Instantiations of the split-BF16 contraction for 1..4 k-steps of 16
// subjects (N up to 64).
#include "bpg_contract_bf16.h"

namespace bpg {
template hipError_t dispatch_bf16_pp<1>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<2>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<3>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<4>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_persist_pp<1>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<2>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<3>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<4>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
}  // namespace bpg
