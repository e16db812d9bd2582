// mt_host.cpp -- MT19937 jump-ahead polynomials (host side).
//
// Generator: M. Matsumoto and T. Nishimura, ACM TOMACS 8 (1998) 3
// (init_genrand, init_by_array, the twist recurrence).  Jump-ahead:
// H. Haramoto et al., "Efficient jump ahead for F2-linear random number
// generators", INFORMS J. Computing 20 (2008) 385: the state moves by the
// transition A; A^e = g(A) with g = x^e mod phi, phi the characteristic
// polynomial of A (degree 19937), so the state e steps ahead is
// sum_i g_i A^i state.  phi is obtained here with the Berlekamp-Massey
// algorithm from 2 * 19937 terms of one bit of the raw sequence.
//
// State convention: a window (x_t, ..., x_{t+623}) of the raw recurrence
// x_{n+624} = x_{n+397} ^ ((x_n^u | x_{n+1}^l) >> 1) ^ (x_{n+1} odd ? a : 0);
// the next output of the window is temper(x_{t+624}).  init_* gives the
// window at t = 0, so stream word w is temper(x_{624+w}).
#include <map>
#include <mutex>
#include <vector>
#include <stdint.h>
#include <string.h>

#include "mt.h"

namespace {

constexpr int MT_N = 624, MT_M = 397, MT_DEG = 19937;
constexpr int PW = (MT_DEG + 1 + 63) / 64;          // words of a reduced poly

uint32_t mt_f(uint32_t x0, uint32_t x1, uint32_t xm) {
  const uint32_t y = (x0 & 0x80000000u) | (x1 & 0x7fffffffu);
  return xm ^ (y >> 1) ^ ((x1 & 1u) ? 0x9908b0dfu : 0u);
}

typedef std::vector<uint64_t> Poly;

int deg_of(const Poly &p) {
  for (int i = (int) p.size() - 1; i >= 0; i--)
    if (p[i]) return i * 64 + 63 - __builtin_clzll(p[i]);
  return -1;
}

bool bit(const Poly &p, int i) { return (p[i >> 6] >> (i & 63)) & 1; }

// p ^= q << s
void xor_shift(Poly &p, const Poly &q, int s) {
  const int ws = s >> 6, bs = s & 63;
  const int need = (int) q.size() + ws + 1;
  if ((int) p.size() < need) p.resize(need, 0);
  for (size_t i = 0; i < q.size(); i++) {
    if (!q[i]) continue;
    p[i + ws] ^= q[i] << bs;
    if (bs) p[i + ws + 1] ^= q[i] >> (64 - bs);
  }
}

struct State {
  Poly phi;                     // characteristic polynomial, degree 19937
  std::map<uint64_t, std::vector<uint32_t>> cache;   // e -> coefficient words
  bool ok = false;
};

State &st() {
  static State s;
  return s;
}
std::mutex g_mu;
std::once_flag g_once;

// r mod phi (in place), deg r < 2 * MT_DEG
void reduce(Poly &r, const Poly &phi) {
  for (int i = deg_of(r); i >= MT_DEG; i--)
    if (bit(r, i)) xor_shift(r, phi, i - MT_DEG);
  r.resize(PW, 0);
}

Poly sqr_mod(const Poly &a, const Poly &phi) {
  Poly r(2 * PW + 2, 0);
  for (int i = 0; i <= deg_of(a); i++)
    if (bit(a, i)) r[(2 * i) >> 6] |= 1ull << ((2 * i) & 63);
  reduce(r, phi);
  return r;
}

Poly mulx_mod(const Poly &a, const Poly &phi) {
  Poly r(PW + 1, 0);
  for (int i = PW - 1; i >= 0; i--) {
    r[i + 1] |= a[i] >> 63;
    r[i] |= a[i] << 1;
  }
  reduce(r, phi);
  return r;
}

void compute_phi(State &s) {
  // 2 * 19937 terms of the top bit of the raw sequence from an arbitrary
  // (non-zero) seed
  const int NT = 2 * MT_DEG;
  std::vector<uint32_t> x(MT_N + NT);
  uint32_t w[MT_N];
  mt_seed_window(5489u, w);
  for (int i = 0; i < MT_N; i++) x[i] = w[i];
  for (int n = 0; n < NT; n++) x[MT_N + n] = mt_f(x[n], x[n + 1], x[n + MT_M]);
  std::vector<uint8_t> seq(NT);
  for (int n = 0; n < NT; n++) seq[n] = (uint8_t) (x[MT_N + n] >> 31);
  // Berlekamp-Massey over GF(2): connection polynomial C, s[n] = sum C_i s[n-i]
  std::vector<uint8_t> C(NT + 1, 0), B(NT + 1, 0), T;
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < NT; n++) {
    uint8_t d = seq[n];
    for (int i = 1; i <= L; i++) d ^= C[i] & seq[n - i];
    if (!d) { m++; continue; }
    if (2 * L <= n) {
      T = C;
      for (int i = 0; i + m <= NT; i++) C[i + m] ^= B[i];
      L = n + 1 - L;
      B = T;
      m = 1;
    }
    else {
      for (int i = 0; i + m <= NT; i++) C[i + m] ^= B[i];
      m++;
    }
  }
  if (L != MT_DEG) return;
  // phi(x) = x^L C(1/x)
  s.phi.assign(PW + 1, 0);
  for (int i = 0; i <= L; i++)
    if (C[i]) s.phi[(L - i) >> 6] |= 1ull << ((L - i) & 63);
  // check: the window jumped by 1000 steps equals 1000 direct steps
  std::vector<uint32_t> g;
  {
    Poly r(PW, 0);
    r[0] = 1;                                // x^0
    for (int k = 0; k < 1000; k++) r = mulx_mod(r, s.phi);
    g.assign(MT_N + 1, 0);
    for (int i = 0; i < MT_DEG; i++)
      if (bit(r, i)) g[i >> 5] |= 1u << (i & 31);
  }
  std::vector<uint32_t> seqw(MT_N + MT_DEG + 1000);
  for (int i = 0; i < MT_N; i++) seqw[i] = w[i];
  for (size_t n = MT_N; n < seqw.size(); n++)
    seqw[n] = mt_f(seqw[n - MT_N], seqw[n - MT_N + 1], seqw[n - MT_N + MT_M]);
  uint32_t acc[MT_N];
  memset(acc, 0, sizeof(acc));
  for (int i = 0; i < MT_DEG; i++)
    if ((g[i >> 5] >> (i & 31)) & 1)
      for (int j = 0; j < MT_N; j++) acc[j] ^= seqw[i + j];
  // words 1..623 of the window are exact; the low 31 bits of word 0 are
  // not part of the state
  bool same = (acc[0] & 0x80000000u) == (seqw[1000] & 0x80000000u);
  for (int j = 1; j < MT_N; j++) same = same && acc[j] == seqw[1000 + j];
  s.ok = same;
}

}  // namespace

const uint32_t *mt_jump_poly(uint64_t e) {
  std::call_once(g_once, [] { compute_phi(st()); });
  State &s = st();
  if (!s.ok) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = s.cache.find(e);
  if (it != s.cache.end()) return it->second.data();
  // x^e mod phi by square-and-multiply-by-x over the bits of e
  Poly r(PW, 0);
  r[0] = 1;
  for (int b = 63; b >= 0; b--) {
    if (deg_of(r) > 0 || r[0] != 1) r = sqr_mod(r, s.phi);
    if ((e >> b) & 1) r = mulx_mod(r, s.phi);
  }
  std::vector<uint32_t> g(MT_JUMP_WORDS, 0);
  for (int i = 0; i < MT_DEG; i++)
    if (bit(r, i)) g[i >> 5] |= 1u << (i & 31);
  auto res = s.cache.emplace(e, std::move(g));
  return res.first->second.data();
}
