// Checkpoint/resume and the reference's configuration files (host code).
//
// Checkpoint: everything the dynamics depends on -- spins (FP32, exact),
// temperatures, cones, the rung assignment, the sweep and exchange counters,
// the seed and the schedule -- plus the energy bookkeeping, the acceptance
// counters and, when present, the learned proposal's weights. Philox is
// counter-based, so no generator state exists: a run resumed from a
// checkpoint continues bit for bit. Little-endian binary:
//
//   "PYROCKP1"  u32 version=2  u32 R  u64 seed  f64 jeff
//   i32 n_metro n_learned n_or pt_every   u32 sweep pt_round
//   i32 L[R]  f64 T[R]  f32 cos_cone[R]  i32 replica_at[R]
//   f64 anchor[R]  f64 dE[R]  u64 accepted[R]  u64 proposed[R]
//   u32 has_mlp  [f32 w[4868] if has_mlp]
//   f64 spin_len2  u32 model (bit 0: exchange matrices, bit 1: rotation
//   after the Metropolis test)  [f32 coup[156] if bit 0]
//   u64 n_floats  f32 spins[n_floats]  u64 FNV-1a of every preceding byte
// Version 1 files (no model block: scalar model, no rotation) still load.
//
// Reference files (SaveConfigurations / ReadConfigurations, Math.cu:618-701):
// one file "<prefix><k>.dat" per temperature rank k (descending T), records
// of 3N values in AoS order (s0x s0y s0z s1x ...; spin i = site*ncell + cell),
// FP64 (or FP32), appended one record per repetition.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "pyro_internal.h"

namespace pyro {

namespace {

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
};

struct Writer {
  FILE *f;
  Fnv sum;
  bool ok = true;
  void put(const void *p, size_t n) {
    if (!ok) return;
    ok = fwrite(p, 1, n, f) == n;
    sum.add(p, n);
  }
  template <typename T> void put(const T &v) { put(&v, sizeof v); }
};

struct Reader {
  FILE *f;
  Fnv sum;
  bool ok = true;
  void get(void *p, size_t n) {
    if (!ok) return;
    ok = fread(p, 1, n, f) == n;
    if (ok) sum.add(p, n);
  }
  template <typename T> T get() {
    T v{};
    get(&v, sizeof v);
    return v;
  }
};

const char kMagic[8] = {'P', 'Y', 'R', 'O', 'C', 'K', 'P', '1'};

}  // namespace

}  // namespace pyro

using namespace pyro;

extern "C" {

int pyro_checkpoint_save(pyro_ctx *ctx, const char *path) {
  if (!ctx || !path) return set_error(PYRO_EINVAL, "pyro_checkpoint_save: null argument");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  const int R = ctx->R;
  std::vector<double> T(R), anchor(R), dE(R);
  std::vector<float> cosc(R), spins(ctx->total_floats);
  std::vector<int> at(R);
  std::vector<uint64_t> acc(R);
  std::vector<uint32_t> ctl(kCtlWords);
  PYRO_HIP(hipMemcpy(T.data(), ctx->d_temps, sizeof(double) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(anchor.data(), ctx->d_anchor, sizeof(double) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(dE.data(), ctx->d_dE, sizeof(double) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(cosc.data(), ctx->d_cosc, sizeof(float) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(at.data(), ctx->d_replica_at, sizeof(int) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(acc.data(), ctx->d_acc, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(ctl.data(), ctx->d_ctl, sizeof(uint32_t) * kCtlWords, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(spins.data(), ctx->d_spins, sizeof(float) * ctx->total_floats,
                     hipMemcpyDeviceToHost));
  std::vector<float> w;
  if (ctx->d_mlp_w && ctx->mlp_ready) {
    w.resize(kMlpParams);
    PYRO_HIP(hipMemcpy(w.data(), ctx->d_mlp_w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  }
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return set_error(PYRO_EIO, "pyro_checkpoint_save: cannot open %s", tmp.c_str());
  Writer wr{f};
  wr.put(kMagic, 8);
  wr.put((uint32_t)2);
  wr.put((uint32_t)R);
  wr.put((uint64_t)ctx->seed);
  wr.put((double)ctx->jeff);
  wr.put((int32_t)ctx->n_metro);
  wr.put((int32_t)ctx->n_learned);
  wr.put((int32_t)ctx->n_or);
  wr.put((int32_t)ctx->pt_every);
  wr.put(ctl[kCtlSweep]);
  wr.put(ctl[kCtlPtRound]);
  for (int r = 0; r < R; ++r) wr.put((int32_t)ctx->h_L[r]);
  wr.put(T.data(), sizeof(double) * R);
  wr.put(cosc.data(), sizeof(float) * R);
  wr.put(at.data(), sizeof(int) * R);
  wr.put(anchor.data(), sizeof(double) * R);
  wr.put(dE.data(), sizeof(double) * R);
  wr.put(acc.data(), sizeof(uint64_t) * R);
  wr.put(ctx->h_proposed.data(), sizeof(uint64_t) * R);
  wr.put((uint32_t)(w.empty() ? 0 : 1));
  if (!w.empty()) wr.put(w.data(), sizeof(float) * kMlpParams);
  wr.put((double)ctx->spin_len2);
  wr.put((uint32_t)((ctx->mat ? 1u : 0u) | (ctx->post_rotate ? 2u : 0u)));
  if (ctx->mat) wr.put(ctx->h_coup, sizeof ctx->h_coup);
  wr.put((uint64_t)ctx->total_floats);
  wr.put(spins.data(), sizeof(float) * ctx->total_floats);
  const uint64_t h = wr.sum.h;
  if (wr.ok) wr.ok = fwrite(&h, sizeof h, 1, f) == 1;
  const bool closed = fclose(f) == 0;
  if (!wr.ok || !closed) {
    std::remove(tmp.c_str());
    return set_error(PYRO_EIO, "pyro_checkpoint_save: write to %s failed", tmp.c_str());
  }
  if (std::rename(tmp.c_str(), path) != 0)
    return set_error(PYRO_EIO, "pyro_checkpoint_save: cannot rename to %s", path);
  return 0;
}

int pyro_checkpoint_load(pyro_ctx *ctx, const char *path) {
  if (!ctx || !path) return set_error(PYRO_EINVAL, "pyro_checkpoint_load: null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return set_error(PYRO_EIO, "pyro_checkpoint_load: cannot open %s", path);
  Reader rd{f};
  char magic[8];
  rd.get(magic, 8);
  const uint32_t version = rd.get<uint32_t>();
  const uint32_t R = rd.get<uint32_t>();
  auto bad = [&](const char *what) {
    fclose(f);
    return set_error(PYRO_EIO, "pyro_checkpoint_load: %s (%s)", what, path);
  };
  if (!rd.ok || std::memcmp(magic, kMagic, 8) != 0) return bad("not a checkpoint");
  if (version != 1 && version != 2) return bad("unknown version");
  if ((int)R != ctx->R) return bad("replica count differs from the ctx");
  const uint64_t seed = rd.get<uint64_t>();
  const double jeff = rd.get<double>();
  const int32_t n_metro = rd.get<int32_t>(), n_learned = rd.get<int32_t>();
  const int32_t n_or = rd.get<int32_t>(), pt_every = rd.get<int32_t>();
  const uint32_t sweep = rd.get<uint32_t>(), pt_round = rd.get<uint32_t>();
  std::vector<int32_t> L(R);
  rd.get(L.data(), sizeof(int32_t) * R);
  if (!rd.ok) return bad("truncated header");
  for (uint32_t r = 0; r < R; ++r)
    if (L[r] != ctx->h_L[r]) return bad("lattice sizes differ from the ctx");
  std::vector<double> T(R), anchor(R), dE(R);
  std::vector<float> cosc(R);
  std::vector<int> at(R);
  std::vector<uint64_t> acc(R), prop(R);
  rd.get(T.data(), sizeof(double) * R);
  rd.get(cosc.data(), sizeof(float) * R);
  rd.get(at.data(), sizeof(int) * R);
  rd.get(anchor.data(), sizeof(double) * R);
  rd.get(dE.data(), sizeof(double) * R);
  rd.get(acc.data(), sizeof(uint64_t) * R);
  rd.get(prop.data(), sizeof(uint64_t) * R);
  const uint32_t has_mlp = rd.get<uint32_t>();
  std::vector<float> w;
  if (has_mlp) {
    w.resize(kMlpParams);
    rd.get(w.data(), sizeof(float) * kMlpParams);
  }
  double spin_len2 = ctx->spin_len2;
  uint32_t model = 0;
  float coup[kCoupFloats] = {};
  if (version >= 2) {
    spin_len2 = rd.get<double>();
    model = rd.get<uint32_t>();
    if (model & 1u) rd.get(coup, sizeof coup);
  }
  const uint64_t n_floats = rd.get<uint64_t>();
  if (!rd.ok) return bad("truncated");
  if (n_floats != ctx->total_floats) return bad("spin count differs from the ctx");
  std::vector<float> spins(n_floats);
  rd.get(spins.data(), sizeof(float) * n_floats);
  const uint64_t expect = rd.sum.h;
  uint64_t h = 0;
  const bool got = rd.ok && fread(&h, sizeof h, 1, f) == 1;
  fclose(f);
  if (!got) return set_error(PYRO_EIO, "pyro_checkpoint_load: truncated (%s)", path);
  if (h != expect) return set_error(PYRO_EIO, "pyro_checkpoint_load: checksum mismatch (%s)", path);
  // Validate the contents before touching the ctx.
  std::vector<int> seen(R, 0);
  for (uint32_t k = 0; k < R; ++k) {
    if (at[k] < 0 || at[k] >= (int)R || seen[at[k]]++)
      return set_error(PYRO_EIO, "pyro_checkpoint_load: rung assignment is not a permutation");
    if (!(T[k] > 0.0) || !(cosc[k] >= -1.0f && cosc[k] < 1.0f))
      return set_error(PYRO_EIO, "pyro_checkpoint_load: bad temperature or cone");
  }
  if (n_metro < 0 || n_or < 0 || pt_every < 0 || n_learned < 0)
    return set_error(PYRO_EIO, "pyro_checkpoint_load: bad schedule");
  if (n_learned > 0 && w.empty())
    return set_error(PYRO_EIO, "pyro_checkpoint_load: learned passes without weights");
  if (model > 3u || !(spin_len2 > 0.0)) return set_error(PYRO_EIO, "pyro_checkpoint_load: bad model block");
  for (float c : coup)
    if (!std::isfinite(c)) return set_error(PYRO_EIO, "pyro_checkpoint_load: coupling not finite");
  if (model != 0u && (ctx->hetero || ctx->exec_mode != 0))
    return set_error(PYRO_ESTATE,
                     "pyro_checkpoint_load: the saved model runs in execution mode 0 on a uniform lattice");

  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (!w.empty()) PYRO_TRY(pyro_mlp_set_weights(ctx, w.data()));
  if ((model & 1u) && !ctx->d_coup)
    PYRO_HIP(hipMalloc((void **)&ctx->d_coup, sizeof(float) * kCoupFloats));
  if (model & 1u) PYRO_HIP(hipMemcpy(ctx->d_coup, coup, sizeof coup, hipMemcpyHostToDevice));
  std::vector<float> beta(R);
  for (uint32_t r = 0; r < R; ++r) beta[r] = (float)(1.0 / T[r]);
  PYRO_HIP(hipMemcpy(ctx->d_spins, spins.data(), sizeof(float) * n_floats, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_temps, T.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_beta, beta.data(), sizeof(float) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_cosc, cosc.data(), sizeof(float) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_replica_at, at.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_anchor, anchor.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_dE, dE.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_acc, acc.data(), sizeof(uint64_t) * R, hipMemcpyHostToDevice));
  PYRO_TRY(launch_set_sweep(ctx, sweep, pt_round));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_sweeps = 0;
  ctx->seed = seed;
  ctx->jeff = (float)jeff;
  ctx->n_metro = n_metro;
  ctx->n_learned = n_learned;
  ctx->n_or = n_or;
  ctx->pt_every = pt_every;
  ctx->h_sweep = sweep;
  ctx->h_pt_round = pt_round;
  ctx->h_proposed = prop;
  ctx->spin_len2 = spin_len2;
  ctx->mat = (model & 1u) != 0;
  ctx->post_rotate = (model & 2u) != 0;
  std::memcpy(ctx->h_coup, coup, sizeof coup);
  return 0;
}

static int config_file_name(char *buf, size_t n, const char *prefix, int k) {
  const int w = snprintf(buf, n, "%s%d.dat", prefix, k);
  return (w > 0 && (size_t)w < n) ? 0 : -1;
}

int pyro_export_configurations(pyro_ctx *ctx, const char *prefix, int append, int bytes_per_value) {
  if (!ctx || !prefix) return set_error(PYRO_EINVAL, "pyro_export_configurations: null argument");
  if (ctx->hetero) return set_error(PYRO_ESTATE, "pyro_export_configurations: needs a uniform lattice size");
  if (bytes_per_value != 4 && bytes_per_value != 8)
    return set_error(PYRO_EINVAL, "pyro_export_configurations: bytes_per_value %d not 4/8", bytes_per_value);
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  const int R = ctx->R;
  const size_t N = (size_t)16 * ctx->ncell;
  std::vector<int> at(R);
  std::vector<float> soa(3 * N);
  PYRO_HIP(hipMemcpy(at.data(), ctx->d_replica_at, sizeof(int) * R, hipMemcpyDeviceToHost));
  std::vector<double> aos8;
  std::vector<float> aos4;
  for (int k = 0; k < R; ++k) {
    const int r = at[k];
    PYRO_HIP(hipMemcpy(soa.data(), ctx->d_spins + ctx->h_off[r], sizeof(float) * 3 * N,
                       hipMemcpyDeviceToHost));
    char name[4096];
    if (config_file_name(name, sizeof name, prefix, k))
      return set_error(PYRO_EINVAL, "pyro_export_configurations: prefix too long");
    FILE *f = fopen(name, append ? "ab" : "wb");
    if (!f) return set_error(PYRO_EIO, "pyro_export_configurations: cannot open %s", name);
    bool ok;
    if (bytes_per_value == 8) {
      aos8.resize(3 * N);
      for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) aos8[3 * i + d] = soa[d * N + i];
      ok = fwrite(aos8.data(), sizeof(double), 3 * N, f) == 3 * N;
    } else {
      aos4.resize(3 * N);
      for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) aos4[3 * i + d] = soa[d * N + i];
      ok = fwrite(aos4.data(), sizeof(float), 3 * N, f) == 3 * N;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return set_error(PYRO_EIO, "pyro_export_configurations: write to %s failed", name);
  }
  return 0;
}

int pyro_import_configurations(pyro_ctx *ctx, const char *prefix, int rep, int bytes_per_value) {
  if (!ctx || !prefix || rep < 0) return set_error(PYRO_EINVAL, "pyro_import_configurations: bad argument");
  if (ctx->hetero) return set_error(PYRO_ESTATE, "pyro_import_configurations: needs a uniform lattice size");
  if (bytes_per_value != 4 && bytes_per_value != 8)
    return set_error(PYRO_EINVAL, "pyro_import_configurations: bytes_per_value %d not 4/8", bytes_per_value);
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  const int R = ctx->R;
  const size_t N = (size_t)16 * ctx->ncell;
  std::vector<int> at(R);
  PYRO_HIP(hipMemcpy(at.data(), ctx->d_replica_at, sizeof(int) * R, hipMemcpyDeviceToHost));
  std::vector<float> all(ctx->total_floats);
  std::vector<unsigned char> raw(3 * N * bytes_per_value);
  for (int k = 0; k < R; ++k) {
    char name[4096];
    if (config_file_name(name, sizeof name, prefix, k))
      return set_error(PYRO_EINVAL, "pyro_import_configurations: prefix too long");
    FILE *f = fopen(name, "rb");
    if (!f) return set_error(PYRO_EIO, "pyro_import_configurations: cannot open %s", name);
    const long off = (long)rep * (long)(3 * N * bytes_per_value);
    const bool ok = fseek(f, off, SEEK_SET) == 0 && fread(raw.data(), 1, raw.size(), f) == raw.size();
    fclose(f);
    if (!ok)
      return set_error(PYRO_EIO, "pyro_import_configurations: record %d not in %s (end of file)", rep, name);
    float *dst = all.data() + ctx->h_off[at[k]];
    for (size_t i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) {
        float v;
        if (bytes_per_value == 8) {
          double x;
          std::memcpy(&x, raw.data() + 8 * (3 * i + d), 8);
          v = (float)x;
        } else {
          std::memcpy(&v, raw.data() + 4 * (3 * i + d), 4);
        }
        if (!std::isfinite(v))
          return set_error(PYRO_EIO, "pyro_import_configurations: non-finite value in %s", name);
        dst[d * N + i] = v;
      }
  }
  return pyro_set_spins(ctx, all.data());
}

}  // extern "C"
