// Shared constants and small device helpers for the batched KEM engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dpkem {

// ---- FrodoKEM-976-SHAKE (FrodoKEM976-GPU/include/params.h:9-18, api.h:7-10)
constexpr int kFrodoN = 976;
constexpr int kFrodoNbar = 8;
constexpr int kFrodoPk = 15632;
constexpr int kFrodoSk = 31296;
constexpr int kFrodoCt = 15744;
constexpr int kFrodoSs = 24;
constexpr int kFrodoKeygenCoins = 64;  // s(24) || seedSE(24) || z(16)
constexpr int kFrodoEncCoins = 24;     // mu
constexpr int kFrodoSeedA = 16;
constexpr int kFrodoSec = 24;          // len(s) = len(seedSE) = len(k) = len(pkh)

// ---- Saber l=3 (SABER-GPU/include/params.h:6-48)
constexpr int kSaberN = 256;
constexpr int kSaberL = 3;
constexpr int kSaberPk = 992;
constexpr int kSaberSk = 2304;
constexpr int kSaberCt = 1088;
constexpr int kSaberSs = 32;
constexpr int kSaberKeygenCoins = 96;  // seedA_raw(32) || seed_s(32) || z(32)
constexpr int kSaberEncCoins = 32;     // m_raw
constexpr int kSaberPolyVecBytes = 1248;
constexpr int kSaberPolyVecCompBytes = 960;
constexpr int kSaberGenMatrixBytes = 3744;  // 3 * 1248
constexpr int kSaberGenSecretBytes = 768;   // 3 * 256

__device__ __forceinline__ uint64_t ld64(const uint8_t* p) { return *reinterpret_cast<const uint64_t*>(p); }
__device__ __forceinline__ void st64(uint8_t* p, uint64_t v) { *reinterpret_cast<uint64_t*>(p) = v; }

__device__ __forceinline__ uint32_t popc(uint32_t x) { return __builtin_popcount(x); }

}  // namespace dpkem
