/*
 * gprt_hip_ext.h — declared extensions of gprt_hip.h with no counterpart in GPRT's host API:
 * width-explicit radix sorts.
 *
 * The reference's sort templates advertise uint32_t buffers (gprt/gprt_host.h:2390, :2423-2429)
 * while its implementation reads 64-bit keys and payloads (gprt/gprt.cpp:8837); gprtBufferSort
 * and gprtBufferSortPayload in gprt_hip.h keep that 64-bit behaviour. These entry points make the
 * width explicit. All are stable, ascending, in place and enqueued on the context's stream (they
 * return at once); `scratch` follows the reference's rule (null = internal allocation; an
 * undersized buffer is resized in place, gprt/gprt_host.h:2380-2382). Results do not depend on
 * gprtSetSortConfig.
 */
#ifndef GPRT_HIP_EXT_H
#define GPRT_HIP_EXT_H

#include "gprt_hip.h"

/* Sorts every element of `keys`: u32 keys (count = size / 4) or u64 keys (count = size / 8). */
GPRT_API void gprtBufferSort32(GPRTContext context, GPRTBuffer keys, GPRTBuffer scratch GPRT_IF_CPP(= nullptr));
GPRT_API void gprtBufferSort64(GPRTContext context, GPRTBuffer keys, GPRTBuffer scratch GPRT_IF_CPP(= nullptr));

/* Key + payload of one width: u32 keys with u32 values (count = keys size / 4), or u64 keys with
 * u64 values (count = keys size / 8). The value buffer must hold at least `count` values. */
GPRT_API void gprtBufferSortPayload32(GPRTContext context, GPRTBuffer keys, GPRTBuffer values,
                                      GPRTBuffer scratch GPRT_IF_CPP(= nullptr));
GPRT_API void gprtBufferSortPayload64(GPRTContext context, GPRTBuffer keys, GPRTBuffer values,
                                      GPRTBuffer scratch GPRT_IF_CPP(= nullptr));

/* Any combination: keyBits = 32 or 64; valueBits = 0 (no payload), 32 or 64; `count` elements
 * (keys need not fill the buffer). */
GPRT_API void gprtBufferSortKV(GPRTContext context, GPRTBuffer keys, GPRTBuffer values, uint32_t count,
                               uint32_t keyBits, uint32_t valueBits, GPRTBuffer scratch);

#endif /* GPRT_HIP_EXT_H */
