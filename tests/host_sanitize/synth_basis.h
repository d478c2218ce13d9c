// Synthetic basis and shell-tuple lists shared by the host-sanitizer
// drivers (test infrastructure).
#pragma once
#include <cmath>
#include <random>
#include <vector>

namespace synth {

const int ENV0 = 20;  // first user slot of env

struct SynthBasis {
  std::vector<int> atm, bas;
  std::vector<double> env;
  int natm = 0, nbas = 0;
};

// natm atoms on a jittered line; every atom carries s, s, p, p, d shells and
// a generally contracted s (two columns), even atoms also an f and a g
inline SynthBasis make_basis(int natm, double omega, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  SynthBasis B;
  B.natm = natm;
  B.env.assign(ENV0, 0.0);
  B.env[8] = omega;
  B.atm.assign(6 * natm, 0);
  for (int a = 0; a < natm; ++a) {
    B.atm[6 * a + 0] = 1 + a % 8;
    B.atm[6 * a + 1] = (int)B.env.size();
    B.env.push_back(1.6 * a + 0.3 * U(rng));
    B.env.push_back(0.4 * U(rng));
    B.env.push_back(0.4 * U(rng));
  }
  // (l, nprim, nctr) per shell on every atom; heavier shells on even atoms
  const int shells[][3] = {{0, 3, 1}, {0, 1, 1}, {1, 2, 1}, {1, 1, 1}, {2, 1, 1}, {0, 2, 2}};
  const int heavy[][3] = {{3, 1, 1}, {4, 1, 1}};
  auto add = [&](int atom, int l, int np, int nc) {
    const int pe = (int)B.env.size();
    for (int i = 0; i < np; ++i) B.env.push_back(0.15 * std::pow(4.0, i) * (1.0 + 0.5 * U(rng)));
    const int pc = (int)B.env.size();
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < np; ++i) B.env.push_back(0.2 + U(rng));
    const int row[8] = {atom, l, np, nc, 0, pe, pc, 0};
    B.bas.insert(B.bas.end(), row, row + 8);
  };
  for (int a = 0; a < natm; ++a) {
    for (auto& s : shells) add(a, s[0], s[1], s[2]);
    if (a % 2 == 0)
      for (auto& s : heavy) add(a, s[0], s[1], s[2]);
  }
  B.nbas = (int)B.bas.size() / 8;
  return B;
}

// unique quartets (i>=j, k>=l, ij>=kl) of the first n shells
inline std::vector<int> unique_quartets(int n) {
  std::vector<int> q;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= i; ++k)
        for (int l = 0; l <= (k == i ? j : k); ++l) q.insert(q.end(), {i, j, k, l});
  return q;
}

inline int nao_sph(const SynthBasis& B) {
  int n = 0;
  for (int s = 0; s < B.nbas; ++s) n += (2 * B.bas[8 * s + 1] + 1) * B.bas[8 * s + 3];
  return n;
}


}  // namespace synth
