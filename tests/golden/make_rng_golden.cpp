// Writes rng_golden.json: known-answer vectors for the random-number layer of
// the transport path, produced by the libraries the reference links against.
//  * pcg32 from the upstream pcg-cpp header (the copy vendored with pyarrow in
//    this image; Abeille pins HunterBelanger/pcg-cpp @ feature/cmake, whose
//    engine code is upstream pcg-cpp),
//  * libstdc++ 11 <random>: generate_canonical via uniform_real_distribution,
//    exponential_distribution and discrete_distribution,
// exactly as include/utils/rng.hpp and include/simulation/particle.hpp call
// them.  Built and run by make_rng_golden.sh; the JSON is the fixture.
#include <pcg_random.hpp>

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using pcg32 = arrow_vendored::pcg32;

static void print_u64_list(const std::vector<unsigned long long>& v) {
  std::printf("[");
  for (size_t i = 0; i < v.size(); i++) std::printf("%s%llu", i ? ", " : "", v[i]);
  std::printf("]");
}

static void print_dbl_list(const std::vector<double>& v) {
  std::printf("[");
  for (size_t i = 0; i < v.size(); i++) std::printf("%s\"%a\"", i ? ", " : "", v[i]);
  std::printf("]");
}

int main() {
  const unsigned long long seed = 19073486328125ULL;
  const unsigned long long stride = 152917ULL;
  const unsigned long long hist[] = {0ULL, 1ULL, 2ULL, 999ULL, 123456789ULL};
  std::printf("{\n  \"seed\": %llu, \"stride\": %llu,\n  \"streams\": [\n", seed, stride);
  for (int hi = 0; hi < 5; hi++) {
    const unsigned long long h = hist[hi];
    pcg32 rng(seed);
    rng.advance(stride * h);
    std::vector<unsigned long long> raw;
    pcg32 r2 = rng;
    for (int i = 0; i < 8; i++) raw.push_back(r2());
    // uniforms as RNG::rand draws them
    std::uniform_real_distribution<double> unit(0., 1.);
    pcg32 r3 = rng;
    std::vector<double> uni;
    for (int i = 0; i < 8; i++) uni.push_back(unit(r3));
    // exponential with lambda = 1.7 (glibc log)
    pcg32 r4 = rng;
    std::vector<double> ex;
    for (int i = 0; i < 8; i++) {
      std::exponential_distribution<double> d(1.7);
      ex.push_back(d(r4));
    }
    // state after advance, to pin Brown's jump
    std::printf("    {\"history\": %llu, \"raw_u32\": ", h);
    print_u64_list(raw);
    std::printf(", \"rand\": ");
    print_dbl_list(uni);
    std::printf(", \"exponential_1.7\": ");
    print_dbl_list(ex);
    std::printf("}%s\n", hi < 4 ? "," : "");
  }
  std::printf("  ],\n");

  // discrete_distribution: results and draw consumption for several weight sets
  std::printf("  \"discrete\": [\n");
  const std::vector<std::vector<double>> wsets = {
      {0.30, 0.15},
      {0.0, 1.0},
      {1.0},
      {0.2, 0.0, 0.5, 0.3},
      {0.1, 0.2, 0.3, 0.4, 0.05, 0.05, 0.9},
  };
  for (size_t w = 0; w < wsets.size(); w++) {
    pcg32 rng(seed);
    rng.advance(stride * 7);
    std::vector<unsigned long long> res;
    for (int i = 0; i < 16; i++) {
      std::discrete_distribution<int> d(wsets[w].begin(), wsets[w].end());
      res.push_back((unsigned long long)d(rng));
    }
    // next raw output tells how many outputs were consumed in total
    unsigned long long next = rng();
    std::printf("    {\"weights\": ");
    print_dbl_list(wsets[w]);
    std::printf(", \"results\": ");
    print_u64_list(res);
    std::printf(", \"next_raw_u32\": %llu}%s\n", next, w + 1 < wsets.size() ? "," : "");
  }
  std::printf("  ],\n");

  // advance(n) must equal n single steps
  {
    pcg32 a(seed), b(seed);
    a.advance(1000003ULL);
    for (int i = 0; i < 1000003; i++) b();
    std::printf("  \"advance_1000003_next\": [%u, %u],\n", (unsigned)a(), (unsigned)b());
  }
  // O'Neill's pcg32 demo: pcg32 rng(42u, 54u) first outputs
  {
    pcg32 demo(42u, 54u);
    std::vector<unsigned long long> v;
    for (int i = 0; i < 6; i++) v.push_back(demo());
    std::printf("  \"demo_42_54\": ");
    print_u64_list(v);
    std::printf("\n}\n");
  }
  return 0;
}
