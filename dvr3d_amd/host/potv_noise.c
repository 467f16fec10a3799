/* Index-free noise of the synthetic ABC surface (potv_example.f90, kind 2):
 * a standard normal number keyed by the IEEE bit patterns of (r1, r2, x), so
 * that the same point always gets the same value whatever grid it lies on.
 *   h  = mix(bits(x) ^ mix(bits(r2) ^ mix(bits(r1) ^ seed)))
 *   u1 = ((h >> 11) + 1) 2^-53,  u2 = (mix(h) >> 11) 2^-53
 *   N  = sqrt(-2 ln u1) cos(2 pi u2)          (Box-Muller, cosine branch)
 * mix is the splitmix64 finaliser (Steele, Lea, Flood 2014), seed 20251203. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t bits(double v) {
  uint64_t b;
  memcpy(&b, &v, sizeof b);
  return b;
}

double dvr_synth_normal(double r1, double r2, double x) {
  const uint64_t seed = 20251203ull;
  uint64_t h = mix64(bits(r1) ^ seed);
  h = mix64(bits(r2) ^ h);
  h = mix64(bits(x) ^ h);
  const uint64_t h2 = mix64(h);
  const double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}
