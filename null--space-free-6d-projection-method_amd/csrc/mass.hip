// Instantiation and launch of the three mass-operator passes for every grid
// size nGrid = 2N (N = 1..6) in FP64 and FP32, with the tile variants the
// autotuner chooses between.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace fame6d {

// Tile variant tables.  Pass A/C: BPW = S-blocks per workgroup (threads = BPW*n^2).
// Pass B: A = consecutive inner positions per workgroup (threads = A*n^2,
// LDS = A*n^3 complex values).  Every entry divides S = n^3.
template <int n, class T>
struct Tiles;
// Mhat passes (three thread groups): BPW3 blocks per workgroup, threads = 3*BPW3*n^2.
// Pass B variant v: A = a[v] inner positions per workgroup, vp[v] of them per thread (2: the
// FP32 form with 16-byte accesses, k_pass_b2).
#define F6_TILES(N_, T_, BPW0, BPW1, A0, A1, B30, B31, VP0, VP1) \
  template <>                                                   \
  struct Tiles<N_, T_> {                                        \
    static constexpr int bpw[2] = {BPW0, BPW1};                 \
    static constexpr int a[2] = {A0, A1};                       \
    static constexpr int bpw3[2] = {B30, B31};                  \
    static constexpr int vp[2] = {VP0, VP1};                    \
  };
F6_TILES(2, double, 8, 8, 8, 8, 8, 8, 1, 1)
F6_TILES(2, float, 8, 8, 8, 8, 8, 8, 1, 1)
F6_TILES(4, double, 16, 8, 16, 32, 8, 4, 1, 1)
F6_TILES(4, float, 16, 8, 16, 32, 8, 4, 1, 1)
F6_TILES(6, double, 8, 6, 8, 24, 4, 2, 1, 1)
F6_TILES(6, float, 8, 6, 8, 24, 4, 2, 1, 1)
F6_TILES(8, double, 4, 8, 8, 4, 2, 1, 1, 1)
F6_TILES(8, float, 4, 8, 8, 16, 2, 4, 1, 1)
F6_TILES(10, double, 2, 1, 4, 2, 1, 2, 1, 1)
F6_TILES(10, float, 2, 1, 4, 8, 1, 2, 1, 2)
F6_TILES(12, double, 2, 2, 4, 2, 1, 1, 1, 1)
F6_TILES(12, float, 2, 2, 6, 8, 1, 2, 1, 2)
#undef F6_TILES

template <class F>
void with_n(int n, F&& f) {
  switch (n) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 10: f(std::integral_constant<int, 10>{}); break;
    case 12: f(std::integral_constant<int, 12>{}); break;
    default: throw Error(FAME6D_EARG, "unsupported grid size nGrid=" + std::to_string(n));
  }
}

// The tile choice of a launch over ngroups groups: launches over several groups take the
// choice tuned on several groups (one tile per workgroup: the grid is already several times
// the device, and the persistent forms tuned on one group only add their loop there).
template <class T>
static const TileChoice& tile_of(const fame6d_mass* m, int ngroups) {
  if constexpr (std::is_same<T, double>::value) return ngroups > 1 ? m->tile64g : m->tile64;
  else return ngroups > 1 ? m->tile32g : m->tile32;
}

// The group count that selects the choice: the Mhat passes (three vectors per group) take the
// several-group choice from two groups on; the plain passes, whose one-group choice is tuned
// on one and two right-hand sides (the pairs around each K_r^{-1}), from three.
static int tile_groups(bool mhat, int ngroups) { return mhat || ngroups > 2 ? ngroups : 1; }

// Workgroups per group of the persistent Mhat passes A and C: per_cu per compute unit, at
// most one per tile.
static int pass_c3p_blocks(const fame6d_mass* m, int tiles, int per_cu) {
  return std::max(1, std::min(tiles, per_cu * m->cus));
}

template <class T>
void launch_pass_a(const fame6d_mass* m, int op, int ngroups, const PassAArgs<T>& a, hipStream_t s) {
  if (ngroups <= 0) return;
  const TileChoice& tc = tile_of<T>(m, tile_groups(op == PRE_MHAT || op == PRE_MHAT_CG, ngroups));
  const int v = tc.a;
  if (op == PRE_MHAT || op == PRE_MHAT_CG) {
    const int v3 = tc.a3;
    const int a3p = tc.a3p, a3s = tc.a3s;
    with_n(m->n, [&](auto nc) {
      constexpr int n = decltype(nc)::value;
      constexpr int S = n * n * n;
      auto go = [&](auto bc) {
        constexpr int BPW = decltype(bc)::value;
        if (a3p > 0) {
          dim3 grid(pass_c3p_blocks(m, S / BPW, a3p), ngroups), block(3 * BPW * n * n);
          if (a3s && op == PRE_MHAT) launch_k(m, k_pass_a3s<T, n, BPW, PRE_MHAT>, grid, block, s, a);
          else if (a3s) launch_k(m, k_pass_a3s<T, n, BPW, PRE_MHAT_CG>, grid, block, s, a);
          else if (op == PRE_MHAT) launch_k(m, k_pass_a3p<T, n, BPW, PRE_MHAT>, grid, block, s, a);
          else launch_k(m, k_pass_a3p<T, n, BPW, PRE_MHAT_CG>, grid, block, s, a);
          return;
        }
        dim3 grid(S / BPW, ngroups), block(3 * BPW * n * n);
        if (op == PRE_MHAT) launch_k(m, k_pass_a3<T, n, BPW, PRE_MHAT>, grid, block, s, a);
        else launch_k(m, k_pass_a3<T, n, BPW, PRE_MHAT_CG>, grid, block, s, a);
      };
      if (v3 == 0) go(std::integral_constant<int, Tiles<n, T>::bpw3[0]>{});
      else go(std::integral_constant<int, Tiles<n, T>::bpw3[1]>{});
    });
    F6_CHECK_LAUNCH();
    return;
  }
  with_n(m->n, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    constexpr int S = n * n * n;
    auto go = [&](auto bc) {
      constexpr int BPW = decltype(bc)::value;
      dim3 grid(S / BPW, ngroups), block(BPW * n * n);
      if (op == PRE_V1H || op == PRE_ZW) {  // FP64 only, coalesced form
        if constexpr (std::is_same<T, double>::value) {
          if (op == PRE_V1H) launch_k(m, k_pass_a_lin<T, n, BPW, PRE_V1H>, grid, block, s, a);
          else launch_k(m, k_pass_a_lin<T, n, BPW, PRE_ZW>, grid, block, s, a);
          return;
        }
        throw Error(FAME6D_EARG, "bad pass-A op");
      }
      if (tc.io == 1) {
        switch (op) {
          case PRE_COPY: launch_k(m, k_pass_a_lin<T, n, BPW, PRE_COPY>, grid, block, s, a); break;
          case PRE_MHAT: launch_k(m, k_pass_a_lin<T, n, BPW, PRE_MHAT>, grid, block, s, a); break;
          case PRE_MHAT_CG: launch_k(m, k_pass_a_lin<T, n, BPW, PRE_MHAT_CG>, grid, block, s, a); break;
          default: throw Error(FAME6D_EARG, "bad pass-A op");
        }
      } else {
        switch (op) {
          case PRE_COPY: launch_k(m, k_pass_a<T, n, BPW, PRE_COPY>, grid, block, s, a); break;
          case PRE_MHAT: launch_k(m, k_pass_a<T, n, BPW, PRE_MHAT>, grid, block, s, a); break;
          case PRE_MHAT_CG: launch_k(m, k_pass_a<T, n, BPW, PRE_MHAT_CG>, grid, block, s, a); break;
          default: throw Error(FAME6D_EARG, "bad pass-A op");
        }
      }
    };
    if (v == 0) go(std::integral_constant<int, Tiles<n, T>::bpw[0]>{});
    else go(std::integral_constant<int, Tiles<n, T>::bpw[1]>{});
  });
  F6_CHECK_LAUNCH();
}

template <class T>
void launch_pass_b(const fame6d_mass* m, int nr, int mode, int ngroups, cpx<T>* t, long long ldt,
                   Skip skip, hipStream_t s, bool inverse_eps, double* partials) {
  if (ngroups <= 0) return;
  const void* e = std::is_same<T, double>::value ? (inverse_eps ? m->ieps64.ptr : m->eps64.ptr)
                                                 : (inverse_eps ? m->ieps32.ptr : m->eps32.ptr);
  PassBArgs<T> a{t, ldt, (const T*)e, skip, partials};
  const TileChoice& tc = tile_of<T>(m, tile_groups(nr == 3, ngroups));
  const int v = nr == 1 ? tc.b1 : tc.b;  // plain applies have their own tile
  with_n(m->n, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    constexpr int S = n * n * n;
    auto go = [&](auto ac, auto vc) {
      constexpr int A = decltype(ac)::value;
      constexpr int VP = decltype(vc)::value;
      dim3 grid(S / A, ngroups * nr);
      if constexpr (VP == 2 && std::is_same<T, float>::value) {
        dim3 block(A / 2 * n * n);
        if (mode == 0 && nr == 1) launch_k(m, k_pass_b2<n, A, 1, 0>, grid, block, s, a);
        else if (mode == 0 && nr == 3) launch_k(m, k_pass_b2<n, A, 3, 0>, grid, block, s, a);
        else if (mode == 1 && nr == 1) launch_k(m, k_pass_b2<n, A, 1, 1>, grid, block, s, a);
        else throw Error(FAME6D_EARG, "bad pass-B configuration");
      } else {
        static_assert(VP == 1, "two positions per thread only in FP32");
        dim3 block(A * n * n);
        if (mode == 0 && nr == 1) launch_k(m, k_pass_b<T, n, A, 1, 0>, grid, block, s, a);
        else if (mode == 0 && nr == 3) launch_k(m, k_pass_b<T, n, A, 3, 0>, grid, block, s, a);
        else if (mode == 1 && nr == 1) launch_k(m, k_pass_b<T, n, A, 1, 1>, grid, block, s, a);
        else throw Error(FAME6D_EARG, "bad pass-B configuration");
      }
    };
    if (v == 0)
      go(std::integral_constant<int, Tiles<n, T>::a[0]>{}, std::integral_constant<int, Tiles<n, T>::vp[0]>{});
    else
      go(std::integral_constant<int, Tiles<n, T>::a[1]>{}, std::integral_constant<int, Tiles<n, T>::vp[1]>{});
  });
  F6_CHECK_LAUNCH();
}

template <class T>
void launch_pass_c(const fame6d_mass* m, int op, int ngroups, const PassCArgs<T>& a, hipStream_t s) {
  if (ngroups <= 0) return;
  const TileChoice& tc = tile_of<T>(m, tile_groups(op == POST_MHAT || op == POST_MHAT_CG, ngroups));
  if (op == POST_MHAT || op == POST_MHAT_CG) {
    const int v3 = tc.c3;
    with_n(m->n, [&](auto nc) {
      constexpr int n = decltype(nc)::value;
      constexpr int S = n * n * n;
      const int c3p = tc.c3p;
      auto go = [&](auto bc) {
        constexpr int BPW = decltype(bc)::value;
        if (c3p > 0) {
          dim3 grid(pass_c3p_blocks(m, S / BPW, c3p), ngroups), block(3 * BPW * n * n);
          if (op == POST_MHAT_CG) launch_k(m, k_pass_c3p<T, n, BPW, true>, grid, block, s, a);
          else launch_k(m, k_pass_c3p<T, n, BPW, false>, grid, block, s, a);
          return;
        }
        dim3 grid(S / BPW, ngroups), block(3 * BPW * n * n);
        if (op == POST_MHAT_CG) launch_k(m, k_pass_c3<T, n, BPW, true>, grid, block, s, a);
        else launch_k(m, k_pass_c3<T, n, BPW, false>, grid, block, s, a);
      };
      if (v3 == 0) go(std::integral_constant<int, Tiles<n, T>::bpw3[0]>{});
      else go(std::integral_constant<int, Tiles<n, T>::bpw3[1]>{});
    });
    F6_CHECK_LAUNCH();
    return;
  }
  const int v = tc.a;
  with_n(m->n, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    constexpr int S = n * n * n;
    auto go = [&](auto bc) {
      constexpr int BPW = decltype(bc)::value;
      dim3 grid(S / BPW, ngroups), block(BPW * n * n);
      const bool lin = tc.io == 1;
      if (op == POST_COPY && lin) launch_k(m, k_pass_c_lin<T, n, BPW, POST_COPY, 1>, grid, block, s, a);
      else if (op == POST_MHAT && lin) launch_k(m, k_pass_c_lin<T, n, BPW, POST_MHAT, 3>, grid, block, s, a);
      else if (op == POST_COPY) launch_k(m, k_pass_c<T, n, BPW, POST_COPY, 1>, grid, block, s, a);
      else if (op == POST_MHAT) launch_k(m, k_pass_c<T, n, BPW, POST_MHAT, 3>, grid, block, s, a);
      else throw Error(FAME6D_EARG, "bad pass-C op");
    };
    if (v == 0) go(std::integral_constant<int, Tiles<n, T>::bpw[0]>{});
    else go(std::integral_constant<int, Tiles<n, T>::bpw[1]>{});
  });
  F6_CHECK_LAUNCH();
}

int pass_c_blocks(const fame6d_mass* m, bool fp32, int ngroups) {
  // workgroups per group of the Mhat pass C (the only pass C that writes partials)
  int out = 0;
  const TileChoice& tc = fp32 ? tile_of<float>(m, ngroups) : tile_of<double>(m, ngroups);
  with_n(m->n, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    constexpr int S = n * n * n;
    const int bpw = fp32 ? Tiles<n, float>::bpw3[tc.c3] : Tiles<n, double>::bpw3[tc.c3];
    out = tc.c3p > 0 ? pass_c3p_blocks(m, S / bpw, tc.c3p) : S / bpw;
  });
  return out;
}

int pass_b_blocks(const fame6d_mass* m, bool fp32, int ngroups) {
  // tiles of pass B per vector (size of its partials per vector)
  int out = 0;
  const int v = fp32 ? tile_of<float>(m, ngroups).b : tile_of<double>(m, ngroups).b;
  with_n(m->n, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    constexpr int S = n * n * n;
    const int A = fp32 ? Tiles<n, float>::a[v] : Tiles<n, double>::a[v];
    out = S / A;
  });
  return out;
}

int num_tile_variants(int n, bool fp32, bool pass_b) {
  (void)n;
  (void)fp32;
  (void)pass_b;
  return 2;
}

void tile_variant_desc(int nn, bool fp32, bool pass_b, int v, int* param, int* threads) {
  with_n(nn, [&](auto nc) {
    constexpr int n = decltype(nc)::value;
    int p, per = 1;  // per: inner positions per thread of pass B
    if (pass_b) {
      p = fp32 ? Tiles<n, float>::a[v] : Tiles<n, double>::a[v];
      per = fp32 ? Tiles<n, float>::vp[v] : Tiles<n, double>::vp[v];
    } else {
      p = fp32 ? Tiles<n, float>::bpw[v] : Tiles<n, double>::bpw[v];
    }
    *param = p;
    *threads = p * n * n / per;
  });
}

template <class T>
DevBuf& work_buf(fame6d_mass* m) {
  if constexpr (std::is_same<T, double>::value) return m->work64;
  else return m->work32;
}

// y = M_s t for the two outputs t of a PRE_V1H or PRE_ZW pass A per group (2 * ngroups
// vectors at y + v*ldy, group g's at v = 2g, 2g + 1).  `a` carries the op's inputs.
void mass_apply_pre(fame6d_mass* m, int op, int ngroups, const PassAArgs<double>& a_in, cpx<double>* y,
                    long long ldy, hipStream_t s) {
  if (ngroups <= 0) return;
  if (op != PRE_V1H && op != PRE_ZW) throw Error(FAME6D_EARG, "bad fused M_s apply");
  const int nrhs = 2 * ngroups;
  DevBuf& w = work_buf<double>(m);
  w.ensure(sizeof(cpx<double>) * (size_t)m->NF * (size_t)nrhs);
  cpx<double>* t = w.as<cpx<double>>();
  PassAArgs<double> a = a_in;
  a.t = t;
  a.ldt = m->NF;
  launch_pass_a<double>(m, op, ngroups, a, s);
  launch_pass_b<double>(m, 1, 0, nrhs, t, m->NF, no_skip(), s);
  PassCArgs<double> c{};
  c.t = t;
  c.ldt = m->NF;
  c.y = y;
  c.ldy = ldy;
  launch_pass_c<double>(m, POST_COPY, nrhs, c, s);
}

template <class T>
void mass_apply(fame6d_mass* m, const cpx<T>* x, long long ldx, cpx<T>* y, long long ldy, int nrhs,
                hipStream_t s) {
  if (nrhs <= 0) return;
  DevBuf& w = work_buf<T>(m);
  w.ensure(sizeof(cpx<T>) * (size_t)m->NF * (size_t)nrhs);
  mass_apply_with(m, x, ldx, y, ldy, nrhs, w.as<cpx<T>>(), s);
}

template <class T>
void mass_apply_with(fame6d_mass* m, const cpx<T>* x, long long ldx, cpx<T>* y, long long ldy, int nrhs,
                     cpx<T>* t, hipStream_t s) {
  if (nrhs <= 0) return;
  PassAArgs<T> a{};
  a.x = x;
  a.ldx = ldx;
  a.t = t;
  a.ldt = m->NF;
  launch_pass_a<T>(m, PRE_COPY, nrhs, a, s);
  launch_pass_b<T>(m, 1, 0, nrhs, t, m->NF, no_skip(), s);
  PassCArgs<T> c{};
  c.t = t;
  c.ldt = m->NF;
  c.y = y;
  c.ldy = ldy;
  launch_pass_c<T>(m, POST_COPY, nrhs, c, s);
}

// ------------------------------------------------------------------ host plumbing
void d2h_sync(fame6d_mass* m, void* host, const void* dev, size_t bytes) {
  void* p = m->pin.take(bytes, m->stream);
  F6_HIP(hipMemcpyAsync(p, dev, bytes, hipMemcpyDeviceToHost, m->stream));
  F6_HIP(hipStreamSynchronize(m->stream));
  std::memcpy(host, p, bytes);
  m->pin.used = 0;
}

void h2d_async(fame6d_mass* m, void* dev, const void* host, size_t bytes) {
  void* p = m->pin.take(bytes, m->stream);
  std::memcpy(p, host, bytes);
  F6_HIP(hipMemcpyAsync(dev, p, bytes, hipMemcpyHostToDevice, m->stream));
}

void KernelTimer::enable(int e) {
  if (e > 0 && ev.empty()) {
    ev.resize(2 * kPool);
    // timing-only events: without the system-scope fence a bracketed kernel does not pay for
    // an L2 writeback at its end that unbracketed launches do not have
    for (auto& x : ev) F6_HIP(hipEventCreateWithFlags(&x, hipEventDisableSystemFence));
    slot.resize(kPool);
    bytes_pending.resize(kPool);
    tag.resize(kPool);
  }
  harvest();
  reset();
  on = e > 0;
  stride = e > 0 ? e : 1;
  armed = fired = false;
}

void KernelTimer::begin(int sl) {
  armed = (seen[sl]++ % stride) == 0;
  fired = false;
  if (!armed) return;
  if (pending >= kPool) harvest();
  slot[pending] = sl;
}

void KernelTimer::end(double b, int tg) {
  if (armed && fired) {
    bytes_pending[pending] = b;
    tag[pending] = tg;
    ++pending;
  }
  armed = fired = false;
}

void KernelTimer::discard_tags(int first, int owner, int threshold) {
  for (int i = std::max(first, 0); i < pending; ++i)
    if (tag[i] >= 0 && tag[i] / kTagSpan == owner && tag[i] % kTagSpan >= threshold) bytes_pending[i] = -1.0;
}

void KernelTimer::harvest() {
  if (pending == 0) return;
  F6_HIP(hipEventSynchronize(ev[2 * pending - 1]));
  for (int i = 0; i < pending; ++i) {
    if (bytes_pending[i] < 0) continue;
    float t = 0.f;
    F6_HIP(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
    count[slot[i]] += 1;
    ms[slot[i]] += t;
    bytes[slot[i]] += bytes_pending[i];
  }
  pending = 0;
}

void KernelTimer::reset() {
  for (int i = 0; i < kSlots; ++i) {
    count[i] = seen[i] = 0;
    ms[i] = bytes[i] = 0;
  }
}

void KernelTimer::destroy() {
  for (auto& x : ev) (void)hipEventDestroy(x);
  ev.clear();
  pending = 0;
  on = false;
}

#define F6_INST(T)                                                                                 \
  template void launch_pass_a<T>(const fame6d_mass*, int, int, const PassAArgs<T>&, hipStream_t);  \
  template void launch_pass_b<T>(const fame6d_mass*, int, int, int, cpx<T>*, long long, Skip,      \
                                 hipStream_t, bool, double*);                                      \
  template void launch_pass_c<T>(const fame6d_mass*, int, int, const PassCArgs<T>&, hipStream_t);  \
  template void mass_apply<T>(fame6d_mass*, const cpx<T>*, long long, cpx<T>*, long long, int,     \
                              hipStream_t);                                                        \
  template void mass_apply_with<T>(fame6d_mass*, const cpx<T>*, long long, cpx<T>*, long long, int, \
                                   cpx<T>*, hipStream_t);                                          \
  template DevBuf& work_buf<T>(fame6d_mass*);
F6_INST(double)
F6_INST(float)
#undef F6_INST

}  // namespace fame6d
