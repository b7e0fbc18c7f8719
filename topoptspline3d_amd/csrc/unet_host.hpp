// unet_host.hpp -- host side of the learned coarse-level correction (no counterpart in the
// reference): the 3-level U-Net, its forward / backward passes built from the kernels of
// kernels_nn.hip, Adam training on a device-resident pair dataset.
//
// Network (channels c0/c1/c2 = 16/32/64, 3x3x3 convolutions, 2 per block):
//   enc0: conv 4->16, conv 16->16            (level grid g0)
//   enc1: conv 16->32 stride 2, conv 32->32  (g1 = ceil(g0 / 2) node-aligned)
//   enc2: conv 32->64 stride 2, conv 64->64  (g2)
//   dec1: upsample enc2 -> g1, concat [up 64 | enc1 32], conv 96->32, conv 32->32
//   dec0: upsample dec1 -> g0, concat [up 32 | enc0 16], conv 48->16, conv 16->16
//   out : conv 16->3 (no ReLU, FP32 output)
// ReLU after every other convolution.  Input channels: scaled residual (3) and nodal
// log-density (1), padded to 8.  Parameters: FP32 master W[tap][cin][cout] then b[cout]
// per layer, BF16 copies for the MFMA kernels.
#pragma once

struct tos_unet;
struct tos_dataset;

namespace tos {

struct NetLayer {
  int cin, cinp, cout, nout, stride;
  bool relu;
  int64_t woff, boff;
  int kpf, kpb;
  uint16_t* wf = nullptr;
  uint16_t* wb = nullptr;
};

struct NetActs {
  int B = 0;
  NGrid g0{}, g1{}, g2{};
  bool grads = false;
  uint16_t *feat, *e0a, *cat0, *e1a, *cat1, *e2a, *e2b, *d1a, *d1b, *d0a, *d0b;
  float* y;
  // gradient tensors and loss inputs (training only)
  uint16_t *gy, *g_d0b, *g_d0a, *g_cat0, *g_d1b, *g_d1a, *g_cat1, *g_e2b, *g_e2a, *g_e1b, *g_e1a, *g_e0b, *g_e0a;
  float* tgt;
  uint8_t* mask;
  float* inv_den;
  std::vector<void*> mem;
};

inline int64_t ngrid_v(const NGrid& g) { return (int64_t)g.X * g.Y * g.Z; }
inline NGrid ngrid_half(const NGrid& g) { return NGrid{(g.X + 1) / 2, (g.Y + 1) / 2, (g.Z + 1) / 2}; }

}  // namespace tos

struct tos_unet {
  int dev = 0;
  hipStream_t stream = nullptr;
  int ch[3] = {16, 32, 64};
  std::vector<tos::NetLayer> layers;
  int64_t nparam = 0;
  float *param = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  int64_t adam_t = 0;
  bool dirty = true;  // BF16 weight copies out of date
  std::vector<tos::NetActs*> acts;
  float* wpart = nullptr;  // weight-gradient partials
  size_t wpart_n = 0;
  float* bpart = nullptr;  // bias-gradient partials
  double* lpart = nullptr; // loss partials
  size_t lpart_n = 0;
  double* dscal = nullptr; // device scalars (loss)
  double* h_scal = nullptr;
  std::vector<void*> allocs;
  std::string err;
};

struct tos_dataset {
  int dev = 0;
  hipStream_t stream = nullptr;
  tos::NGrid g{};
  int64_t V = 0;
  int cap = 0, n = 0;
  uint16_t* feat = nullptr;  // cap x V x 8 BF16
  float* tgt = nullptr;      // cap x V x 3
  uint8_t* mask = nullptr;   // cap x V
  float* inv_den = nullptr;  // cap
  std::vector<void*> allocs;
  std::string err;
};

namespace {

constexpr int kWgradGroups = 64;
constexpr int kBiasGroups = 64;

template <class T>
T* nalloc(std::vector<void*>& list, hipStream_t s, size_t n) {
  return (T*)raw_alloc(list, s, n * sizeof(T));
}

// layer table of the 16/32/64 U-Net
void unet_layout(tos_unet* N) {
  const int c0 = N->ch[0], c1 = N->ch[1], c2 = N->ch[2];
  struct D {
    int cin, cout, stride;
    bool relu;
  };
  const D d[11] = {{4, c0, 1, true},       {c0, c0, 1, true}, {c0, c1, 2, true}, {c1, c1, 1, true},
                   {c1, c2, 2, true},      {c2, c2, 1, true}, {c2 + c1, c1, 1, true}, {c1, c1, 1, true},
                   {c1 + c0, c0, 1, true}, {c0, c0, 1, true}, {c0, 3, 1, false}};
  int64_t off = 0;
  N->layers.clear();
  for (int i = 0; i < 11; ++i) {
    tos::NetLayer L{};
    L.cin = d[i].cin;
    L.cinp = (d[i].cin + 7) / 8 * 8;
    L.cout = d[i].cout;
    L.nout = (d[i].cout + 15) / 16 * 16;
    L.stride = d[i].stride;
    L.relu = d[i].relu;
    L.woff = off;
    off += 27LL * L.cin * L.cout;
    L.boff = off;
    off += L.cout;
    L.kpf = (27 * L.cinp + 31) / 32 * 32;
    L.kpb = (27 * L.nout + 31) / 32 * 32;
    N->layers.push_back(L);
  }
  N->nparam = off;
}

// deterministic He initialisation (splitmix64 + Box-Muller), biases 0
void unet_init_params(tos_unet* N, uint64_t seed, std::vector<float>& h) {
  h.assign(N->nparam, 0.f);
  uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ULL;
  auto next = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (size_t li = 0; li < N->layers.size(); ++li) {
    const tos::NetLayer& L = N->layers[li];
    const double sd = std::sqrt(2.0 / (27.0 * L.cin)) * (L.relu ? 1.0 : 0.1);
    const int64_t n = 27LL * L.cin * L.cout;
    for (int64_t i = 0; i < n; ++i) {
      const double u1 = unif(), u2 = unif();
      h[L.woff + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  }
}

tos::NetActs* unet_acts(tos_unet* N, int B, tos::NGrid g0, bool grads) {
  for (tos::NetActs* A : N->acts)
    if (A->B == B && A->g0.X == g0.X && A->g0.Y == g0.Y && A->g0.Z == g0.Z && (A->grads || !grads)) return A;
  tos::NetActs* A = new tos::NetActs();
  A->B = B;
  A->g0 = g0;
  A->g1 = tos::ngrid_half(g0);
  A->g2 = tos::ngrid_half(A->g1);
  A->grads = grads;
  const int c0 = N->ch[0], c1 = N->ch[1], c2 = N->ch[2];
  const int64_t V0 = B * tos::ngrid_v(A->g0), V1 = B * tos::ngrid_v(A->g1), V2 = B * tos::ngrid_v(A->g2);
  auto bf = [&](int64_t n) { return nalloc<uint16_t>(A->mem, N->stream, n); };
  A->feat = bf(V0 * 8);
  A->e0a = bf(V0 * c0);
  A->cat0 = bf(V0 * (c1 + c0));
  A->e1a = bf(V1 * c1);
  A->cat1 = bf(V1 * (c2 + c1));
  A->e2a = bf(V2 * c2);
  A->e2b = bf(V2 * c2);
  A->d1a = bf(V1 * c1);
  A->d1b = bf(V1 * c1);
  A->d0a = bf(V0 * c0);
  A->d0b = bf(V0 * c0);
  A->y = nalloc<float>(A->mem, N->stream, V0 * 3);
  if (grads) {
    A->gy = bf(V0 * 16);
    A->g_d0b = bf(V0 * c0);
    A->g_d0a = bf(V0 * c0);
    A->g_cat0 = bf(V0 * (c1 + c0));
    A->g_d1b = bf(V1 * c1);
    A->g_d1a = bf(V1 * c1);
    A->g_cat1 = bf(V1 * (c2 + c1));
    A->g_e2b = bf(V2 * c2);
    A->g_e2a = bf(V2 * c2);
    A->g_e1b = bf(V1 * c1);
    A->g_e1a = bf(V1 * c1);
    A->g_e0b = bf(V0 * c0);
    A->g_e0a = bf(V0 * c0);
    A->tgt = nalloc<float>(A->mem, N->stream, V0 * 3);
    A->mask = nalloc<uint8_t>(A->mem, N->stream, V0);
    A->inv_den = nalloc<float>(A->mem, N->stream, B);
  }
  HIPCHK(hipStreamSynchronize(N->stream));
  N->acts.push_back(A);
  return A;
}

void unet_free_acts(tos_unet* N) {
  for (tos::NetActs* A : N->acts) {
    for (void* p : A->mem) (void)hipFree(p);
    delete A;
  }
  N->acts.clear();
}

void unet_prep(tos_unet* N, hipStream_t st) {
  if (!N->dirty) return;
  for (size_t li = 0; li < N->layers.size(); ++li) {
    tos::NetLayer& L = N->layers[li];
    const bool bwd = li > 0;
    const int64_t n = (int64_t)L.nout * L.kpf + (bwd ? (int64_t)L.cin * L.kpb : 0);
    hipLaunchKernelGGL(k_prep_weights, dim3(nblk(n)), dim3(256), 0, st, (const float*)(N->param + L.woff), L.cin,
                       L.cout, L.cinp, L.nout, L.kpf, L.wf, L.cin, L.kpb, bwd ? L.wb : (uint16_t*)nullptr);
  }
  LAUNCH_CHECK();
  N->dirty = false;
}

// ---- convolution dispatch -------------------------------------------------------------
#define TOS_CONV_CASE(CI, NO, MD)                                                  \
  if (cinp == CI && nout == NO && mode == MD) {                                    \
    hipLaunchKernelGGL((k_conv<CI, NO, MD>), dim3(blocks), dim3(256), 0, st, a);   \
    return;                                                                        \
  }
void conv_launch(int cinp, int nout, int mode, const tos::ConvArgs& a, hipStream_t st) {
  const int64_t tiles = (int64_t)a.B * a.go.Z * a.go.Y * ((a.go.X + 15) / 16);
  const unsigned blocks = (unsigned)((tiles + 3) / 4);
  if (blocks == 0) return;
  TOS_CONV_CASE(8, 16, 0)
  TOS_CONV_CASE(16, 16, 0)
  TOS_CONV_CASE(16, 32, 1)
  TOS_CONV_CASE(32, 32, 0)
  TOS_CONV_CASE(32, 64, 1)
  TOS_CONV_CASE(64, 64, 0)
  TOS_CONV_CASE(96, 32, 0)
  TOS_CONV_CASE(48, 16, 0)
  TOS_CONV_CASE(16, 16, 2)
  TOS_CONV_CASE(16, 48, 2)
  TOS_CONV_CASE(32, 32, 2)
  TOS_CONV_CASE(32, 96, 2)
  TOS_CONV_CASE(64, 64, 2)
  TOS_CONV_CASE(64, 32, 3)
  TOS_CONV_CASE(32, 16, 3)
  throw std::runtime_error("conv: unsupported shape " + std::to_string(cinp) + "x" + std::to_string(nout) +
                           " mode " + std::to_string(mode));
}
#undef TOS_CONV_CASE

#define TOS_WGRAD_CASE(CI, NO, S)                                                                      \
  if (cinp == CI && nout == NO && stride == S) {                                                       \
    hipLaunchKernelGGL((k_conv_wgrad<CI, NO, S>), dim3(27, kWgradGroups), dim3(256), 0, st, x, xcs, xco, gi, dy, \
                       go, B, part);                                                                   \
    return;                                                                                            \
  }
void wgrad_launch(int cinp, int nout, int stride, const uint16_t* x, int xcs, int xco, tos::NGrid gi,
                  const uint16_t* dy, tos::NGrid go, int B, float* part, hipStream_t st) {
  TOS_WGRAD_CASE(8, 16, 1)
  TOS_WGRAD_CASE(16, 16, 1)
  TOS_WGRAD_CASE(16, 32, 2)
  TOS_WGRAD_CASE(32, 32, 1)
  TOS_WGRAD_CASE(32, 64, 2)
  TOS_WGRAD_CASE(64, 64, 1)
  TOS_WGRAD_CASE(96, 32, 1)
  TOS_WGRAD_CASE(48, 16, 1)
  throw std::runtime_error("wgrad: unsupported shape");
}
#undef TOS_WGRAD_CASE

tos::ConvArgs conv_args(int B, const uint16_t* x, int xcs, int xco, tos::NGrid gi, const uint16_t* w,
                        tos::NGrid go) {
  tos::ConvArgs a{};
  a.x = x; a.xcs = xcs; a.xco = xco; a.gi = gi; a.w = w; a.go = go; a.B = B;
  return a;
}

// forward: layer li, input (x, xcs, xco) on gi -> BF16 (y, ycs, yco) or FP32 y32 on go
void layer_fwd(tos_unet* N, int li, int B, const uint16_t* x, int xcs, int xco, tos::NGrid gi, uint16_t* y, int ycs,
               int yco, float* y32, tos::NGrid go, hipStream_t st) {
  const tos::NetLayer& L = N->layers[li];
  tos::ConvArgs a = conv_args(B, x, xcs, xco, gi, L.wf, go);
  a.bias = N->param + L.boff;
  a.y = y; a.ycs = ycs; a.yco = yco;
  a.y32 = y32; a.nstore = L.cout;
  a.relu = L.relu;
  conv_launch(L.cinp, L.nout, L.stride == 1 ? 0 : 1, a, st);
}

void unet_forward(tos_unet* N, tos::NetActs* A, hipStream_t st) {
  unet_prep(N, st);
  const int B = A->B, c0 = N->ch[0], c1 = N->ch[1], c2 = N->ch[2];
  const int k0 = c1 + c0, k1 = c2 + c1;
  const tos::NGrid g0 = A->g0, g1 = A->g1, g2 = A->g2;
  layer_fwd(N, 0, B, A->feat, 8, 0, g0, A->e0a, c0, 0, nullptr, g0, st);
  layer_fwd(N, 1, B, A->e0a, c0, 0, g0, A->cat0, k0, c1, nullptr, g0, st);
  layer_fwd(N, 2, B, A->cat0, k0, c1, g0, A->e1a, c1, 0, nullptr, g1, st);
  layer_fwd(N, 3, B, A->e1a, c1, 0, g1, A->cat1, k1, c2, nullptr, g1, st);
  layer_fwd(N, 4, B, A->cat1, k1, c2, g1, A->e2a, c2, 0, nullptr, g2, st);
  layer_fwd(N, 5, B, A->e2a, c2, 0, g2, A->e2b, c2, 0, nullptr, g2, st);
  hipLaunchKernelGGL(k_upsample, dim3(nblk(B * tos::ngrid_v(g1) * (c2 / 8))), dim3(256), 0, st,
                     (const uint16_t*)A->e2b, c2, g2, A->cat1, k1, 0, g1, B);
  layer_fwd(N, 6, B, A->cat1, k1, 0, g1, A->d1a, c1, 0, nullptr, g1, st);
  layer_fwd(N, 7, B, A->d1a, c1, 0, g1, A->d1b, c1, 0, nullptr, g1, st);
  hipLaunchKernelGGL(k_upsample, dim3(nblk(B * tos::ngrid_v(g0) * (c1 / 8))), dim3(256), 0, st,
                     (const uint16_t*)A->d1b, c1, g1, A->cat0, k0, 0, g0, B);
  layer_fwd(N, 8, B, A->cat0, k0, 0, g0, A->d0a, c0, 0, nullptr, g0, st);
  layer_fwd(N, 9, B, A->d0a, c0, 0, g0, A->d0b, c0, 0, nullptr, g0, st);
  layer_fwd(N, 10, B, A->d0b, c0, 0, g0, nullptr, 0, 0, A->y, g0, st);
  LAUNCH_CHECK();
}

// data gradient of layer li: gradient dy (contiguous nout channels) on the layer's output
// grid go -> dx (dxcs, dxco) on its input grid gi, plus `add`, times the ReLU mask
void layer_dgrad(tos_unet* N, int li, int B, const uint16_t* dy, tos::NGrid go, uint16_t* dx, int dxcs, int dxco,
                 tos::NGrid gi, const uint16_t* mask, int mcs, int mco, const uint16_t* add, int acs, int aco,
                 hipStream_t st) {
  const tos::NetLayer& L = N->layers[li];
  tos::ConvArgs a = conv_args(B, dy, L.nout, 0, go, L.wb, gi);
  a.y = dx; a.ycs = dxcs; a.yco = dxco;
  a.mask = mask; a.mcs = mcs; a.mco = mco;
  a.add = add; a.acs = acs; a.aco = aco;
  conv_launch(L.nout, L.cin, L.stride == 1 ? 2 : 3, a, st);
}

// weight and bias gradients of layer li into N->grad
void layer_wgrad(tos_unet* N, int li, int B, const uint16_t* x, int xcs, int xco, tos::NGrid gi, const uint16_t* dy,
                 tos::NGrid go, hipStream_t st) {
  const tos::NetLayer& L = N->layers[li];
  wgrad_launch(L.cinp, L.nout, L.stride, x, xcs, xco, gi, dy, go, B, N->wpart, st);
  const int ci16 = (L.cinp + 15) / 16 * 16;
  const int64_t nw = 27LL * L.cin * L.cout;
  hipLaunchKernelGGL(k_wgrad_fold, dim3(nblk(nw)), dim3(256), 0, st, (const float*)N->wpart, kWgradGroups, ci16,
                     L.nout, L.cin, L.cout, N->grad + L.woff);
  hipLaunchKernelGGL(k_bias_grad, dim3(kBiasGroups), dim3(256), 0, st, dy, (int64_t)B * tos::ngrid_v(go), L.nout,
                     N->bpart);
  hipLaunchKernelGGL(k_bias_fold, dim3(1), dim3(64), 0, st, (const float*)N->bpart, kBiasGroups, L.nout, L.cout,
                     N->grad + L.boff);
}

// gradient of the loss w.r.t. all parameters, from A->gy (the loss kernel's output)
void unet_backward(tos_unet* N, tos::NetActs* A, hipStream_t st) {
  const int B = A->B, c0 = N->ch[0], c1 = N->ch[1], c2 = N->ch[2];
  const int k0 = c1 + c0, k1 = c2 + c1;
  const tos::NGrid g0 = A->g0, g1 = A->g1, g2 = A->g2;
  // out
  layer_dgrad(N, 10, B, A->gy, g0, A->g_d0b, c0, 0, g0, A->d0b, c0, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 10, B, A->d0b, c0, 0, g0, A->gy, g0, st);
  // dec0
  layer_dgrad(N, 9, B, A->g_d0b, g0, A->g_d0a, c0, 0, g0, A->d0a, c0, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 9, B, A->d0a, c0, 0, g0, A->g_d0b, g0, st);
  layer_dgrad(N, 8, B, A->g_d0a, g0, A->g_cat0, k0, 0, g0, nullptr, 0, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 8, B, A->cat0, k0, 0, g0, A->g_d0a, g0, st);
  hipLaunchKernelGGL(k_upsample_adj, dim3(nblk(B * tos::ngrid_v(g1) * (c1 / 8))), dim3(256), 0, st,
                     (const uint16_t*)A->g_cat0, k0, 0, g0, A->g_d1b, c1, g1, (const uint16_t*)A->d1b, B);
  // dec1
  layer_dgrad(N, 7, B, A->g_d1b, g1, A->g_d1a, c1, 0, g1, A->d1a, c1, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 7, B, A->d1a, c1, 0, g1, A->g_d1b, g1, st);
  layer_dgrad(N, 6, B, A->g_d1a, g1, A->g_cat1, k1, 0, g1, nullptr, 0, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 6, B, A->cat1, k1, 0, g1, A->g_d1a, g1, st);
  hipLaunchKernelGGL(k_upsample_adj, dim3(nblk(B * tos::ngrid_v(g2) * (c2 / 8))), dim3(256), 0, st,
                     (const uint16_t*)A->g_cat1, k1, 0, g1, A->g_e2b, c2, g2, (const uint16_t*)A->e2b, B);
  // enc2
  layer_dgrad(N, 5, B, A->g_e2b, g2, A->g_e2a, c2, 0, g2, A->e2a, c2, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 5, B, A->e2a, c2, 0, g2, A->g_e2b, g2, st);
  layer_dgrad(N, 4, B, A->g_e2a, g2, A->g_e1b, c1, 0, g1, A->cat1, k1, c2, A->g_cat1, k1, c2, st);
  layer_wgrad(N, 4, B, A->cat1, k1, c2, g1, A->g_e2a, g2, st);
  // enc1
  layer_dgrad(N, 3, B, A->g_e1b, g1, A->g_e1a, c1, 0, g1, A->e1a, c1, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 3, B, A->e1a, c1, 0, g1, A->g_e1b, g1, st);
  layer_dgrad(N, 2, B, A->g_e1a, g1, A->g_e0b, c0, 0, g0, A->cat0, k0, c1, A->g_cat0, k0, c1, st);
  layer_wgrad(N, 2, B, A->cat0, k0, c1, g0, A->g_e1a, g1, st);
  // enc0
  layer_dgrad(N, 1, B, A->g_e0b, g0, A->g_e0a, c0, 0, g0, A->e0a, c0, 0, nullptr, 0, 0, st);
  layer_wgrad(N, 1, B, A->e0a, c0, 0, g0, A->g_e0b, g0, st);
  layer_wgrad(N, 0, B, A->feat, 8, 0, g0, A->g_e0a, g0, st);
  LAUNCH_CHECK();
}

// loss of the batch in A (targets, masks, inv_den loaded) -> A->gy, N->dscal[0]
void unet_loss(tos_unet* N, tos::NetActs* A, hipStream_t st) {
  const int64_t V = tos::ngrid_v(A->g0), n = (int64_t)A->B * V;
  const unsigned nb = nblk(n);
  if (N->lpart_n < nb) {
    N->lpart = nalloc<double>(N->allocs, N->stream, nb);
    N->lpart_n = nb;
    HIPCHK(hipStreamSynchronize(N->stream));
  }
  hipLaunchKernelGGL(k_nn_loss, dim3(nb), dim3(256), 0, st, (const float*)A->y, (const float*)A->tgt,
                     (const uint8_t*)A->mask, (const float*)A->inv_den, V, A->B, A->gy, N->lpart);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)N->lpart, (int)nb, N->dscal);
  LAUNCH_CHECK();
}

void unet_adam(tos_unet* N, double lr, hipStream_t st) {
  N->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = (float)(1.0 - std::pow(0.9, (double)N->adam_t));
  const float bc2 = (float)(1.0 - std::pow(0.999, (double)N->adam_t));
  hipLaunchKernelGGL(k_adam, dim3(nblk(N->nparam)), dim3(256), 0, st, N->param, (const float*)N->grad, N->am, N->av,
                     N->nparam, (float)lr, b1, b2, 1e-8f, bc1, bc2);
  LAUNCH_CHECK();
  N->dirty = true;
}

// per-sample 1 / sum m |t|^2 into A->inv_den (targets and masks already in A)
void batch_inv_den(tos_unet* N, tos::NetActs* A, hipStream_t st) {
  const int64_t V = tos::ngrid_v(A->g0);
  const int nb = 256;
  if (N->lpart_n < (size_t)nb) {
    N->lpart = nalloc<double>(N->allocs, N->stream, nb);
    N->lpart_n = nb;
    HIPCHK(hipStreamSynchronize(N->stream));
  }
  for (int b = 0; b < A->B; ++b) {
    hipLaunchKernelGGL(k_target_norm, dim3(nb), dim3(256), 0, st, (const float*)(A->tgt + (int64_t)b * V * 3),
                       (const uint8_t*)(A->mask + (int64_t)b * V), V, N->lpart);
    hipLaunchKernelGGL(k_inv_den, dim3(1), dim3(256), 0, st, (const double*)N->lpart, nb, A->inv_den + b);
  }
  LAUNCH_CHECK();
}

}  // namespace
