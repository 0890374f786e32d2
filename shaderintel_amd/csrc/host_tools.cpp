// Host-side scene preparation of the reference's MPM class, in float64 like the JavaScript it
// restates: the material table row (src/physics/mpm.js:170-196), shape sampling with the
// xorshift32 generator (mpm.js:683-714) and the record packing of MPM.add (mpm.js:213-227).
// No device work happens here.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/mpm.h"

namespace {

// xorshift32 as in mpm.js:684-685; next() returns a double in [0, 1).
class Xs32 {
 public:
  explicit Xs32(uint32_t seed) : s_(seed ? seed : 1u) {}
  double next() {
    s_ ^= s_ << 13;
    s_ ^= s_ >> 17;
    s_ ^= s_ << 5;
    return (double)s_ / 4294967296.0;
  }

 private:
  uint32_t s_;
};

constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" {

void mpm_material_row(int32_t kind, double density, double stiffness, double poisson, double viscosity, double hardening,
                      double compression, double stretch, double friction_deg, int32_t ppc_axis, float out12[12]) {
  const double kCMax = 0.4;        // mpm.js:48
  const double kMaxViscosity = 4;  // mpm.js:35
  const double vol = 1.0 / std::pow((double)ppc_axis, 3.0);
  const double mass = density * vol;
  const double rho = mass / vol;
  const double c = kCMax * std::min(1.0, std::max(0.01, stiffness));
  const double nu = std::min(0.49, std::max(0.0, poisson));
  const double p_modulus = rho * c * c;
  const double mu = p_modulus * (1 - 2 * nu) / (2 * (1 - nu));
  const double lambda = p_modulus * nu / (1 - nu);
  const double phi = (friction_deg * kPi) / 180;
  const double sinp = std::sin(phi);
  float row[12] = {0};
  const uint32_t k = (uint32_t)kind;
  std::memcpy(&row[0], &k, 4);
  row[1] = (float)mass;
  row[2] = (float)(kind == 0 ? rho * c * c : mu);
  row[3] = (float)lambda;
  row[4] = (float)(std::min(kMaxViscosity, std::max(0.0, viscosity)) * 0.05 * rho);
  row[5] = (float)hardening;
  row[6] = (float)compression;
  row[7] = (float)stretch;
  row[8] = (float)(std::sqrt(2.0 / 3.0) * (2 * sinp) / (3 - sinp));
  std::memcpy(out12, row, sizeof row);
}

int64_t mpm_sample_shape(int32_t shape, const double* prm, double spacing, double jitter, uint32_t seed, float* out,
                         int64_t capacity) {
  if (!prm || !(spacing > 0)) return -1;
  double lo[3], hi[3], ctr[3] = {0, 0, 0}, radius = 0;
  switch (shape) {
    case 0:
      for (int a = 0; a < 3; a++) { lo[a] = prm[a]; hi[a] = prm[3 + a]; }
      break;
    case 1:
      for (int a = 0; a < 3; a++) ctr[a] = prm[a];
      radius = prm[3];
      for (int a = 0; a < 3; a++) { lo[a] = ctr[a] - radius; hi[a] = ctr[a] + radius; }
      break;
    case 2: {
      for (int a = 0; a < 3; a++) ctr[a] = prm[a];
      radius = prm[3];
      const double h = prm[4];
      lo[0] = ctr[0] - radius; lo[1] = ctr[1] - h; lo[2] = ctr[2] - radius;
      hi[0] = ctr[0] + radius; hi[1] = ctr[1] + h; hi[2] = ctr[2] + radius;
      break;
    }
    default:
      return -1;
  }
  int64_t cells[3];
  for (int a = 0; a < 3; a++) cells[a] = std::max<int64_t>(1, (int64_t)std::floor((hi[a] - lo[a]) / spacing));
  Xs32 rng(seed);
  int64_t count = 0;
  for (int64_t z = 0; z < cells[2]; z++)
    for (int64_t y = 0; y < cells[1]; y++)
      for (int64_t x = 0; x < cells[0]; x++) {
        const double px = lo[0] + ((double)x + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        const double py = lo[1] + ((double)y + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        const double pz = lo[2] + ((double)z + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        bool inside = true;
        if (shape == 1)
          inside = std::pow(px - ctr[0], 2) + std::pow(py - ctr[1], 2) + std::pow(pz - ctr[2], 2) <= radius * radius;
        else if (shape == 2)
          inside = std::pow(px - ctr[0], 2) + std::pow(pz - ctr[2], 2) <= radius * radius;
        if (!inside) continue;
        if (out && count < capacity) {
          out[3 * count] = (float)px;
          out[3 * count + 1] = (float)py;
          out[3 * count + 2] = (float)pz;
        }
        count++;
      }
  return count;
}

void mpm_pack_particles(const float* pts, int64_t n, const double origin[3], double dx, const int32_t dims[3],
                        double bound, uint32_t material, const double vel[3], double s, double noise_sigma,
                        uint32_t noise_seed, float* rec) {
  Xs32 rng(noise_seed);
  for (int64_t k = 0; k < n; k++) {
    float r[32] = {0};
    for (int a = 0; a < 3; a++) {
      const double x = ((double)pts[3 * k + a] - origin[a]) / dx;
      r[a] = (float)std::min(((double)dims[a] - bound) - 1e-3, std::max(bound + 1e-3, x));
    }
    std::memcpy(&r[3], &material, 4);
    double v[3] = {vel[0], vel[1], vel[2]};
    if (noise_sigma > 0) {
      double z[4];
      for (int j = 0; j < 2; j++) {  // Box-Muller: two uniform pairs, four normals, three used
        const double u1 = 1.0 - rng.next();
        const double u2 = rng.next();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        z[2 * j] = rad * std::cos(2 * kPi * u2);
        z[2 * j + 1] = rad * std::sin(2 * kPi * u2);
      }
      for (int a = 0; a < 3; a++) v[a] = v[a] + noise_sigma * z[a];
    }
    r[4] = (float)(v[0] * s);
    r[5] = (float)(v[1] * s);
    r[6] = (float)(v[2] * s);
    r[7] = 1.0f;                           // J
    r[20] = 1.0f; r[25] = 1.0f; r[30] = 1.0f;  // F = I
    std::memcpy(rec + 32 * k, r, sizeof r);
  }
}

}  // extern "C"
