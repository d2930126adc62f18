// Build-time generator of the sinf/cosf exception table used by rl_trig.h.
//
// Enumerates every float x with |x| <= 64 (both signs, ~2.2e9 values), computes
// the double-precision sin/cos of rl_trig.h, rounds to nearest, and compares
// with the C library's sinf/cosf.  Inputs where the library returned the other
// neighbour are recorded per 4096-value bucket; `delta` is the largest
// normalised distance to the rounding midpoint seen among them.
//
//   gen_trig_table OUT.bin            write the table
//   gen_trig_table --verify IN.bin    re-enumerate and check that
//                                      round_like_libc() == sinf/cosf everywhere
//
// Build: g++ -O2 -fopenmp -ffp-contract=off (no -ffast-math).
#include "rl_trig.h"

#include <gnu/libc-version.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

using namespace rltrig;

struct FileHeader {
  char magic[8];          // "RLTRIG01"
  uint32_t nbuckets;      // RLT_NBUCKETS
  uint32_t abs_limit;     // RLT_ABS_LIMIT
  uint32_t nkeys_sin;
  uint32_t nkeys_cos;
  double delta_sin;
  double delta_cos;
  char libc_version[32];  // gnu_get_libc_version() of the generating host
};

static inline uint32_t bucket_first_bits(uint32_t b) {
  uint32_t sign = (b >> RLT_SIGN_SHIFT) & 1u;
  uint32_t ab = (b & ((1u << RLT_SIGN_SHIFT) - 1u)) << RLT_BUCKET_SHIFT;
  return ab | (sign << 31);
}

static int generate(const char* path) {
  const uint32_t nb = RLT_NBUCKETS;
  std::vector<uint32_t> cnt_s(nb, 0), cnt_c(nb, 0);
  int nthreads = omp_get_max_threads();
  std::vector<std::vector<uint16_t>> ks(nthreads), kc(nthreads);
  std::vector<double> ds(nthreads, 0.0), dc(nthreads, 0.0);
  long bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    int tid = omp_get_thread_num();
#pragma omp for schedule(static)
    for (long bl = 0; bl < (long)nb; ++bl) {
      uint32_t b = (uint32_t)bl;
      uint32_t first = bucket_first_bits(b);
      if ((first & 0x7fffffffu) > RLT_ABS_LIMIT) continue;
      for (uint32_t k = 0; k < (1u << RLT_BUCKET_SHIFT); ++k) {
        uint32_t xb = first | k;
        if ((xb & 0x7fffffffu) > RLT_ABS_LIMIT) break;
        float x = u2f(xb);
        double sd, cd;
        sincos_d(x, &sd, &cd);
        float so, co;
        double sdist, cdist;
        float sf = neighbours(sd, &so, &sdist);
        float cf = neighbours(cd, &co, &cdist);
        float gs = sinf(x), gc = cosf(x);
        if (f2u(gs) != f2u(sf)) {
          if (f2u(gs) != f2u(so)) {
            ++bad;
          } else {
            ks[tid].push_back((uint16_t)k);
            cnt_s[b]++;
            if (sdist > ds[tid]) ds[tid] = sdist;
          }
        }
        if (f2u(gc) != f2u(cf)) {
          if (f2u(gc) != f2u(co)) {
            ++bad;
          } else {
            kc[tid].push_back((uint16_t)k);
            cnt_c[b]++;
            if (cdist > dc[tid]) dc[tid] = cdist;
          }
        }
      }
    }
  }
  if (bad) {
    fprintf(stderr, "gen_trig_table: %ld inputs where libc differs by more than one neighbour\n", bad);
    return 2;
  }
  std::vector<uint32_t> off_s(nb + 1, 0), off_c(nb + 1, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    off_s[b + 1] = off_s[b] + cnt_s[b];
    off_c[b + 1] = off_c[b] + cnt_c[b];
  }
  std::vector<uint16_t> keys_s, keys_c;
  double dmax_s = 0, dmax_c = 0;
  for (int t = 0; t < nthreads; ++t) {  // static schedule: thread order == bucket order
    keys_s.insert(keys_s.end(), ks[t].begin(), ks[t].end());
    keys_c.insert(keys_c.end(), kc[t].begin(), kc[t].end());
    if (ds[t] > dmax_s) dmax_s = ds[t];
    if (dc[t] > dmax_c) dmax_c = dc[t];
  }
  FileHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "RLTRIG01", 8);
  h.nbuckets = nb;
  h.abs_limit = RLT_ABS_LIMIT;
  h.nkeys_sin = (uint32_t)keys_s.size();
  h.nkeys_cos = (uint32_t)keys_c.size();
  // The probe threshold only has to cover every listed input; widen it by a
  // hair so a value exactly at the maximum is probed.
  h.delta_sin = dmax_s * (1.0 + 1e-9) + 1e-12;
  h.delta_cos = dmax_c * (1.0 + 1e-9) + 1e-12;
  snprintf(h.libc_version, sizeof(h.libc_version), "%s", gnu_get_libc_version());
  FILE* fp = fopen(path, "wb");
  if (!fp) {
    perror(path);
    return 1;
  }
  fwrite(&h, sizeof(h), 1, fp);
  fwrite(off_s.data(), 4, nb + 1, fp);
  fwrite(keys_s.data(), 2, keys_s.size(), fp);
  fwrite(off_c.data(), 4, nb + 1, fp);
  fwrite(keys_c.data(), 2, keys_c.size(), fp);
  fclose(fp);
  printf("gen_trig_table: %u sin / %u cos exceptions, delta %.6f / %.6f, libc %s\n", h.nkeys_sin,
         h.nkeys_cos, h.delta_sin, h.delta_cos, h.libc_version);
  return 0;
}

static int verify(const char* path) {
  FILE* fp = fopen(path, "rb");
  if (!fp) {
    perror(path);
    return 1;
  }
  FileHeader h;
  if (fread(&h, sizeof(h), 1, fp) != 1 || memcmp(h.magic, "RLTRIG01", 8) != 0) {
    fprintf(stderr, "bad table file\n");
    return 1;
  }
  const uint32_t nb = h.nbuckets;
  std::vector<uint32_t> off_s(nb + 1), off_c(nb + 1);
  std::vector<uint16_t> keys_s(h.nkeys_sin), keys_c(h.nkeys_cos);
  size_t ok = fread(off_s.data(), 4, nb + 1, fp) == nb + 1 &&
              fread(keys_s.data(), 2, h.nkeys_sin, fp) == h.nkeys_sin &&
              fread(off_c.data(), 4, nb + 1, fp) == nb + 1 &&
              fread(keys_c.data(), 2, h.nkeys_cos, fp) == h.nkeys_cos;
  fclose(fp);
  if (!ok) {
    fprintf(stderr, "truncated table file\n");
    return 1;
  }
  Table ts{off_s.data(), keys_s.data(), h.delta_sin};
  Table tc{off_c.data(), keys_c.data(), h.delta_cos};
  // the flat bitmaps the device uses, built exactly as the library builds them
  int es = table_min_exponent(off_s.data(), keys_s.data());
  int ec = table_min_exponent(off_c.data(), keys_c.data());
  std::vector<uint8_t> bms(bitmap_bytes(es), 0), bmc(bitmap_bytes(ec), 0);
  table_to_bitmap(off_s.data(), keys_s.data(), es, bms.data());
  table_to_bitmap(off_c.data(), keys_c.data(), ec, bmc.data());
  Bitmap bs{bms.data(), es, h.delta_sin};
  Bitmap bc{bmc.data(), ec, h.delta_cos};
  long bad = 0, bad_bm = 0, total = 0;
#pragma omp parallel for schedule(dynamic, 1 << 16) reduction(+ : bad, bad_bm, total)
  for (long u = 0; u <= (long)RLT_ABS_LIMIT; ++u) {
    for (int sg = 0; sg < 2; ++sg) {
      float x = u2f((uint32_t)u | (sg ? 0x80000000u : 0u));
      float s, c, s2, c2;
      sincosf_libc(x, ts, tc, &s, &c);
      sincosf_libc_bm(x, bs, bc, &s2, &c2);
      float gs = sinf(x), gc = cosf(x);
      if (f2u(s) != f2u(gs)) ++bad;
      if (f2u(c) != f2u(gc)) ++bad;
      if (f2u(s2) != f2u(gs)) ++bad_bm;
      if (f2u(c2) != f2u(gc)) ++bad_bm;
      ++total;
    }
  }
  printf("verify: %ld inputs, %ld mismatches (table), %ld mismatches (bitmap, e_min %d/%d) "
         "(libc %s, table from %s)\n", total, bad, bad_bm, es, ec, gnu_get_libc_version(),
         h.libc_version);
  return (bad || bad_bm) ? 3 : 0;
}

int main(int argc, char** argv) {
  if (argc == 3 && strcmp(argv[1], "--verify") == 0) return verify(argv[2]);
  if (argc == 2) return generate(argv[1]);
  fprintf(stderr, "usage: gen_trig_table OUT.bin | --verify IN.bin\n");
  return 1;
}
