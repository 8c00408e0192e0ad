// Host-side triangle planning of the contraction (no device code): the
// reference's triangle order and the shell-pair cover that decides which
// pair's products read each triangle.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace psx {

// -------------------------------------------------------- triangle list ----
// The triangles (i >= j >= l, l >= i - j, shells Ncut//step .. Nmax) in the
// reference's loop order (pyspectrum.py:415-417), as flat (i, j, l) triples.
inline void triangle_list(int step, int ncut, int nmax, std::vector<int>* ijl) {
  const int j0 = ncut / step;
  ijl->clear();
  for (int i = j0; i <= nmax; ++i)
    for (int j = j0; j <= i; ++j)
      for (int l = std::max(i - j, j0); l <= j; ++l) ijl->insert(ijl->end(), {i, j, l});
}

// ------------------------------------------------------- contract plan ----
// Which shell pair reads each triangle.  B is symmetric in its three shells,
// so a triangle (i >= j >= l) is Sum_x P D with P = d_a d_b for any of its
// pairs (i, j), (i, l), (j, l) and D the remaining shell.  Forming a pair's
// products is the costly part of the contraction, so the pairs are a greedy
// set cover (repeatedly the pair that reads most uncovered triangles, ties to
// the smallest key); each triangle then moves to the chosen pair whose
// (pair, 16-column tile) cell reads most triangles, which keeps the number
// of 16x16 output blocks down.  At Nmax = 40, step 3 this is 335 pairs for
// the 440 pairs (i, j) of the reference loop (the greedy cover alone: 339).
// Pair key = a * (nmax + 1) + b, a >= b.  Deterministic.
inline void triangle_cover(const std::vector<int>& ijl, int nmax, int j0, std::vector<int>* tpair,
                           std::vector<int>* tshell) {
  const int ntri = (int)(ijl.size() / 3), nsh = nmax + 1, nkeys = nsh * nsh;
  // the distinct (pair, remaining shell) readings of triangle t
  auto readings = [&](int t, int (&key)[3], int (&rem)[3]) {
    const int i = ijl[3 * t], j = ijl[3 * t + 1], l = ijl[3 * t + 2];
    int n = 0;
    auto add = [&](int a, int b, int c) {
      const int k = a * nsh + b;
      for (int m = 0; m < n; ++m)
        if (key[m] == k) return;
      key[n] = k; rem[n] = c; ++n;
    };
    add(i, j, l); add(i, l, j); add(j, l, i);
    return n;
  };
  int key[3], rem[3];
  std::vector<int> cnt(nkeys, 0);
  std::vector<std::vector<int>> tris_of(nkeys);
  for (int t = 0; t < ntri; ++t) {
    const int n = readings(t, key, rem);
    for (int m = 0; m < n; ++m) { cnt[key[m]]++; tris_of[key[m]].push_back(t); }
  }
  std::vector<char> covered(ntri, 0), chosen(nkeys, 0);
  tpair->assign(ntri, -1);
  tshell->assign(ntri, -1);
  for (;;) {
    int best = -1;
    for (int k = 0; k < nkeys; ++k)
      if (cnt[k] > 0 && (best < 0 || cnt[k] > cnt[best])) best = k;
    if (best < 0) break;
    chosen[best] = 1;
    for (int t : tris_of[best]) {
      if (covered[t]) continue;
      covered[t] = 1;
      const int n = readings(t, key, rem);
      for (int m = 0; m < n; ++m) cnt[key[m]]--;
    }
  }
  // Improve the greedy cover by a fixed number of local-search steps: drop 3
  // to 25 chosen pairs at random, cover the triangles left uncovered again
  // greedily (ties broken at random), prune every pair whose triangles are
  // all read by other chosen pairs (fewest triangles first), and keep the
  // result when it has no more pairs than before.  The generator has a fixed
  // seed, so the plan is the same on every run.  (Nmax 40, step 3: 339 ->
  // 335 pairs, 22 -> 21 row tiles.)
  if (ntri > 0) {
    std::vector<int> cc(ntri, 0);  // chosen readings of each triangle
    for (int k = 0; k < nkeys; ++k)
      if (chosen[k])
        for (int t : tris_of[k]) cc[t]++;
    uint64_t rs = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {  // xorshift64*
      rs ^= rs >> 12; rs ^= rs << 25; rs ^= rs >> 27;
      return (rs * 0x2545F4914F6CDD1Dull) >> 33;
    };
    std::vector<char> ch;
    std::vector<int> c2, gain(nkeys), keys, order;
    std::vector<uint32_t> tie(nkeys);
    int nbest = 0;
    for (int k = 0; k < nkeys; ++k) nbest += chosen[k];
    for (int it = 0; it < 400; ++it) {
      ch = chosen; c2 = cc;
      keys.clear();
      for (int k = 0; k < nkeys; ++k)
        if (ch[k]) keys.push_back(k);
      const int kdrop = std::min((int)keys.size(), 3 + (int)(rnd() % 23));
      for (int d = 0; d < kdrop; ++d) {  // partial Fisher-Yates
        const int r = d + (int)(rnd() % (keys.size() - d));
        std::swap(keys[d], keys[r]);
        ch[keys[d]] = 0;
        for (int t : tris_of[keys[d]]) c2[t]--;
      }
      std::fill(gain.begin(), gain.end(), 0);
      int nunc = 0;
      for (int t = 0; t < ntri; ++t)
        if (c2[t] == 0) {
          ++nunc;
          const int n = readings(t, key, rem);
          for (int m = 0; m < n; ++m) gain[key[m]]++;
        }
      while (nunc > 0) {
        int best = -1, nties = 0;
        for (int k = 0; k < nkeys; ++k) {
          if (ch[k] || gain[k] == 0) continue;
          if (best < 0 || gain[k] > gain[best]) { best = k; nties = 1; }
          else if (gain[k] == gain[best] && rnd() % (uint64_t)(++nties) == 0) best = k;
        }
        ch[best] = 1;
        for (int t : tris_of[best]) {
          if (c2[t]++ == 0) {
            --nunc;
            const int n = readings(t, key, rem);
            for (int m = 0; m < n; ++m) gain[key[m]]--;
          }
        }
      }
      order.clear();
      for (int k = 0; k < nkeys; ++k)
        if (ch[k]) { order.push_back(k); tie[k] = (uint32_t)rnd(); }
      std::sort(order.begin(), order.end(), [&](int a, int b) {
        const size_t na = tris_of[a].size(), nb = tris_of[b].size();
        return na != nb ? na < nb : (tie[a] != tie[b] ? tie[a] < tie[b] : a < b);
      });
      int nch = (int)order.size();
      for (int k : order) {
        bool spare = true;
        for (int t : tris_of[k]) spare = spare && c2[t] >= 2;
        if (!spare) continue;
        ch[k] = 0; --nch;
        for (int t : tris_of[k]) c2[t]--;
      }
      if (nch <= nbest) { chosen.swap(ch); cc.swap(c2); nbest = nch; }
    }
  }
  // each triangle read by its first chosen pair
  for (int t = 0; t < ntri; ++t) {
    const int n = readings(t, key, rem);
    for (int m = 0; m < n; ++m)
      if (chosen[key[m]]) { (*tpair)[t] = key[m]; (*tshell)[t] = rem[m]; break; }
  }
  // move triangles to the fullest (pair, column tile) cell among the chosen pairs
  auto cell = [&](int k, int s) { return k * 4 + (s - j0) / 16; };
  std::vector<int> fill(nkeys * 4, 0);
  for (int t = 0; t < ntri; ++t) fill[cell((*tpair)[t], (*tshell)[t])]++;
  for (int pass = 0; pass < 5; ++pass) {
    for (int t = 0; t < ntri; ++t) {
      const int cur = cell((*tpair)[t], (*tshell)[t]);
      const int n = readings(t, key, rem);
      int bk = (*tpair)[t], bs = (*tshell)[t], bscore = fill[cur] - 1;
      for (int m = 0; m < n; ++m) {
        if (!chosen[key[m]] || key[m] == (*tpair)[t]) continue;
        const int score = fill[cell(key[m], rem[m])];
        if (score > bscore) { bk = key[m]; bs = rem[m]; bscore = score; }
      }
      if (bk != (*tpair)[t]) {
        fill[cur]--;
        fill[cell(bk, bs)]++;
        (*tpair)[t] = bk; (*tshell)[t] = bs;
      }
    }
  }
}

}  // namespace psx
