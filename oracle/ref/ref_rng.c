/* ref_rng.c -- test infrastructure for ref_harness.F90: the project's
 * counter-based splitmix64 draw (stream 0 = ghost events, stream 1 =
 * terminal-bead capture draws), which replaces the absent IFPORT drandm in
 * the harness's restatement of main.F90:1508-1515 exactly as it does in the
 * CPU oracle and on the device.  splitmix64 as published by S. Vigna
 * (constants 0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9, 0x94D049BB133111EB,
 * shifts 30/27/31); the 53 high bits of an output give a double in [0, 1). */
#include <stdint.h>

static uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double ref_u01(uint64_t seed, int32_t stream, uint64_t ctr) {
    const uint64_t key = mix(seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(stream + 1)));
    return (double)(mix(key + ctr) >> 11) * (1.0 / 9007199254740992.0);
}
