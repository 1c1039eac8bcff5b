// Instantiations of the split-BF16 contraction for 5..8 k-steps of 16
// subjects (N up to 128).
#include "bpg_contract_bf16.h"

namespace bpg {
template hipError_t dispatch_bf16_pp<5>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<6>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<7>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_bf16_pp<8>(const ContractArgs&, const Bf16Launch&, int, bool, hipStream_t, int*);
template hipError_t dispatch_persist_pp<5>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<6>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<7>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
template hipError_t dispatch_persist_pp<8>(const BatchClass&, int, bool, int, bool, int, hipStream_t, int*);
}  // namespace bpg
