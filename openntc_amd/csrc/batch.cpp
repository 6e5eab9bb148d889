// Several independent texture sets trained together. Each set is an
// openntc::Context of its own (own latents, decoder, targets and schedule;
// the reference trains one set per Context, libopenntc.cpp:536-879) and all
// share one decoder shape. A training step is one launch per stage for all
// sets: the fused step, the ordered decoder-gradient reduction with Adam, the
// BF16 table refresh and the grid update (per set as Context::Train does it:
// deferred zero-gradient steps, a catch-up of the step's footprint cells
// beside the step before it and the gradient step in them after the fused
// step, or the tile-aware dense update where the tiles cover much of the
// set). Every per-step
// parameter sits in device tables indexed by a device step counter, so the
// step is captured once as a hipGraph and the same graph is replayed for
// every step. Per set, the arithmetic is that of Context::Train's fused path.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "context_impl.h"

using namespace openntc;

namespace
{
struct Tuned
{
  int pass, grid;
};
std::mutex g_tune_mu;
std::map<std::string, Tuned> g_tune_cache;  // decoder shape + set sizes -> launch shape

// Splits `budget` workgroups over sets with work[s] units each: one per set,
// then one at a time to the set with the most work per workgroup, never
// more than cap[s]. Writes the n + 1 offsets.
void split_greedy(const std::vector<int64_t>& work, const std::vector<int64_t>& cap, int budget, int32_t* first)
{
  const int n = (int)work.size();
  std::vector<int64_t> cnt(n, 1);
  for (int left = budget - n; left > 0; --left)
  {
    int best = -1;
    double bw = 0.0;
    for (int s = 0; s < n; ++s)
    {
      if (cnt[s] >= cap[s]) continue;
      const double w = (double)work[s] / (double)cnt[s];
      if (w > bw)
      {
        bw = w;
        best = s;
      }
    }
    if (best < 0) break;
    ++cnt[best];
  }
  first[0] = 0;
  for (int s = 0; s < n; ++s) first[s + 1] = first[s] + (int32_t)cnt[s];
}
} // namespace

struct ntc_batch
{
  std::vector<ntc_context*> sets;
  int n = 0, steps = 0, step_i = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // the deferred sets' catch-up, beside each step
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  static constexpr int kSideCtr = 16;  // the side stream's step counter: step_ctr[16], its own 64 bytes
  int H = 0, F = 0, O = 0, act = 0, max_blocks = 0;
  ntc_feature_desc fd = {};  // level-0 description of set 0: the kernels' shape
  int pass = 0, fused_grid = 0, grid_grid = 0;
  DeviceBuffer fparams, wg_first, dsets, dsteps, psets, trecs, blk_first, step_ctr;
  DeviceBuffer crecs, cfirst, hrecs, hfirst;  // deferred grid steps: catch-up and gradient step
  int lazy_cgrid = 0, lazy_hgrid = 0, lazy_slots = 0;
  bool any_eager = false;
  std::vector<int64_t> step_texels;  // all sets' texels of each step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t texels_done = 0;
  int tuned = 0;  // 1: tuned now, 2: from the cache, 0: given
  float tune_ms[4] = {-1.f, -1.f, -1.f, -1.f};
  int tune_pass[4] = {}, tune_grid[4] = {};

  Context::Impl& impl(int s) const { return *sets[s]->ctx.GetImpl(); }

  ~ntc_batch()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (DeviceBuffer* b : {&fparams, &wg_first, &dsets, &dsteps, &psets, &trecs, &blk_first, &step_ctr, &crecs,
                            &cfirst, &hrecs, &hfirst})
      b->Free();
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // Fused-launch records [steps][n] and workgroup offsets [steps][n + 1] for
  // a pass size and a workgroup budget; returns the largest grid needed.
  int FusedTables(int pass_texels, int budget, std::vector<uint8_t>& rec, std::vector<int32_t>& first) const
  {
    const size_t rb = ntc_fused_params_bytes();
    rec.assign((size_t)steps * n * rb, 0);
    first.assign((size_t)steps * (n + 1), 0);
    std::vector<int64_t> work(n), cap(n);
    int grid = 1;
    for (int t = 0; t < steps; ++t)
    {
      for (int s = 0; s < n; ++s)
      {
        const Context::Impl& I = impl(s);
        const StepPlan& st = I.plan[t];
        const int T = st.tile_dim;
        const int64_t B = (int64_t)st.n_tiles * T * T;
        work[s] = (B + pass_texels - 1) / pass_texels;
        cap[s] = std::min<int64_t>(work[s], I.max_blocks);
        const int lvl = st.level;
        ntc_fused_args a = {};
        a.feat = I.FeatureDesc(lvl);
        a.tiles = {I.origins_table.as<int32_t>() + (int64_t)t * 16, st.n_tiles, T, I.mip_dim[st.lod],
                   (float)st.lod / (float)(I.mip_count - 1)};
        a.noise = {I.seed, (uint32_t)t, lvl, st.grid_update, I.g0_delta, I.g1_delta};
        a.hidden_dim = H;
        a.out_dim = O;
        a.activation = act;
        a.g0 = I.g0[lvl].as<float>();
        a.g1 = I.g1[lvl].as<float>();
        a.weights = I.wtab.as<uint16_t>();
        a.targets = I.package[st.lod].as<uint8_t>();
        a.loss_scale = -2.0f / ((float)O * (float)B);
        a.scatter = st.grid_update;
        a.dg0 = I.dg0[lvl].as<float>();
        a.dg1 = I.dg1[lvl].as<float>();
        a.dw_partials = I.dw_part.as<float>();
        a.sse_partials = I.sse_part.as<double>();
        a.blocks = 1;
        if (ntc_fused_params_pack(&a, &rec[((size_t)t * n + s) * rb]) != NTC_OK) return -1;
      }
      int32_t* f = &first[(size_t)t * (n + 1)];
      split_greedy(work, cap, budget, f);
      grid = std::max(grid, (int)f[n]);
    }
    return grid;
  }

  Result Upload(DeviceBuffer& b, const void* src, size_t bytes)
  {
    if (b.bytes < bytes)
    {
      b.Free();
      if (b.Alloc(bytes) != hipSuccess) return Result::AllocationFailure;
    }
    if (hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return Result::CudaFailure;
    return Result::Success;
  }

  // The stage launches of one step on the main stream (the captured
  // sequence); they read the step counter step_ctr[0] and advance it.
  int LaunchStep()
  {
    int st = ntc_fused_train_sets(&fd, H, act, pass, fparams.ptr, wg_first.as<int32_t>(), n, step_ctr.as<int32_t>(),
                                  fused_grid, stream);
    if (st == NTC_OK)
      st = ntc_decoder_update_sets(H, F, O, static_cast<const ntc_dec_set*>(dsets.ptr),
                                   static_cast<const ntc_dec_step*>(dsteps.ptr), n, step_ctr.as<int32_t>(), stream);
    if (st == NTC_OK)
      st = ntc_pack_decoder_bf16_sets(&fd, H, O, static_cast<const ntc_pack_set*>(psets.ptr), n, stream);
    if (st == NTC_OK && lazy_hgrid > 0)
      st = ntc_grid_lazy_hot_sets(hrecs.ptr, hfirst.as<int32_t>(), n, step_ctr.as<int32_t>(), lazy_hgrid, lazy_slots,
                                  stream);
    if (st == NTC_OK && any_eager)
      st = ntc_grid_update_tiles_sets(trecs.ptr, blk_first.as<int32_t>(), n, step_ctr.as<int32_t>(), grid_grid,
                                      stream);
    if (st == NTC_OK) st = ntc_step_advance(step_ctr.as<int32_t>(), stream);
    return st;
  }

  // The deferred sets' catch-up for the step after step_i, on the side
  // stream beside step_i's launches: it waits for what the main stream held
  // before them (the previous step's gradient step), reads its own counter
  // step_ctr[kSideCtr] (it runs while the main stream advances step_ctr[0])
  // and the next step's launches wait for it. Its records leave out
  // everything step_i reads or writes.
  int LaunchCatchup()
  {
    int32_t* ctr = step_ctr.as<int32_t>() + kSideCtr;
    hipError_t e = hipStreamWaitEvent(side, fork_ev, 0);
    if (e != hipSuccess) return (int)e;
    int st = ntc_grid_lazy_catchup_sets(crecs.ptr, cfirst.as<int32_t>(), n, ctr, lazy_cgrid, lazy_slots, side);
    if (st == NTC_OK) st = ntc_step_advance(ctr, side);
    if (st == NTC_OK && (e = hipEventRecord(join_ev, side)) != hipSuccess) st = (int)e;
    return st;
  }

  // Times the fused launch of the first steps for each (pass size, budget)
  // candidate and keeps the fastest. The fused kernel adds latent gradients
  // into the sets' gradient buffers, so those are cleared afterwards (they
  // are zero between steps); its other outputs are per-step scratch.
  Result Tune(const std::string& key)
  {
    int32_t sizes[2] = {};
    const int ns = ntc_fused_pass_sizes(&fd, H, act, sizes);
    if (ns < 1) return Result::InvalidState;
    const int budgets[2] = {max_blocks, 2 * max_blocks};
    const int k = std::min(steps, 8);
    DeviceBuffer p, w, c;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return Result::CudaFailure;
    Result res = Result::Success;
    float best = 1e30f;
    int ci = 0;
    const int32_t zero = 0;
    for (int i = 0; i < ns && res == Result::Success; ++i)
      for (int j = 0; j < 2 && res == Result::Success; ++j, ++ci)
      {
        std::vector<uint8_t> rec;
        std::vector<int32_t> first;
        const int grid = FusedTables(sizes[i], budgets[j], rec, first);
        if (grid < 1)
        {
          res = Result::InvalidState;
          break;
        }
        if ((res = Upload(p, rec.data(), rec.size())) != Result::Success) break;
        if ((res = Upload(w, first.data(), first.size() * sizeof(int32_t))) != Result::Success) break;
        if ((res = Upload(c, &zero, sizeof(zero))) != Result::Success) break;
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && res == Result::Success; ++rep)  // the first repetition warms up
        {
          if ((res = Upload(c, &zero, sizeof(zero))) != Result::Success) break;
          if (hipEventRecord(e0, stream) != hipSuccess) res = Result::CudaFailure;
          for (int t = 0; t < k && res == Result::Success; ++t)
            if (ntc_fused_train_sets(&fd, H, act, sizes[i], p.ptr, w.as<int32_t>(), n, c.as<int32_t>(), grid,
                                     stream) != NTC_OK ||
                ntc_step_advance(c.as<int32_t>(), stream) != NTC_OK)
              res = Result::CudaFailure;
          if (res == Result::Success &&
              (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
               hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            res = Result::CudaFailure;
        }
        tune_ms[ci] = ms / (float)k;
        tune_pass[ci] = sizes[i];
        tune_grid[ci] = budgets[j];
        if (res == Result::Success && ms < best)
        {
          best = ms;
          pass = sizes[i];
          fused_grid = budgets[j];
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p.Free();
    w.Free();
    c.Free();
    for (int s = 0; s < n && res == Result::Success; ++s)
    {
      Context::Impl& I = impl(s);
      for (int l = 0; l < I.level_count; ++l)
        if (ntc_fill_zero(I.dg0[l].ptr, I.dg0[l].bytes, stream) != NTC_OK ||
            ntc_fill_zero(I.dg1[l].ptr, I.dg1[l].bytes, stream) != NTC_OK)
          res = Result::CudaFailure;
    }
    if (res == Result::Success && hipStreamSynchronize(stream) != hipSuccess) res = Result::CudaFailure;
    if (res == Result::Success)
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune_cache[key] = Tuned{pass, fused_grid};
    }
    return res;
  }

  Result Setup(int pass_texels, int budget)
  {
    Context::Impl& I0 = impl(0);
    H = I0.hidden_dim;
    F = I0.feature_dim;
    O = I0.out_dim;
    act = I0.ActCode();
    fd = I0.FeatureDesc(0);
    steps = I0.batch_count;
    max_blocks = I0.max_blocks;
    for (int s = 0; s < n; ++s)
    {
      const Context::Impl& I = impl(s);
      const ntc_feature_desc f = I.FeatureDesc(0);
      if (!I.manifest_loaded || I.phase != TrainPhase::TrainInProgress || I.path != TrainPath::Fused ||
          I.batch_i != 0 || I.batch_count != steps || I.hidden_dim != H || I.feature_dim != F || I.out_dim != O ||
          I.ActCode() != act || f.g0_channels != fd.g0_channels || f.g1_channels != fd.g1_channels ||
          f.g0_mode != fd.g0_mode || f.pe_count != fd.pe_count || f.pe_first_period != fd.pe_first_period ||
          f.lod_slot != fd.lod_slot || f.g0_channels % 4 || f.g1_channels % 4 || I.grids_per_batch > 16 ||
          I.max_blocks != max_blocks)
        return Result::InvalidState;
    }
    if (!ntc_fused_supported(&fd, H, act)) return Result::InvalidState;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&join_ev, hipEventDisableTiming) != hipSuccess)
      return Result::CudaFailure;

    // launch shape: given, cached for these sets, or tuned now
    std::string key = std::to_string(H) + "/" + std::to_string(F) + "/" + std::to_string(O) + "/" +
                      std::to_string(act) + "/" + std::to_string(fd.g0_channels) + "/" +
                      std::to_string(fd.g1_channels) + "/" + std::to_string(fd.g0_mode) + ":";
    for (int s = 0; s < n; ++s)
      key += std::to_string(impl(s).mip_dim[0]) + "x" + std::to_string(impl(s).grids_per_batch) + ",";
    if (pass_texels > 0)
    {
      pass = pass_texels;
      fused_grid = budget > 0 ? budget : max_blocks;
    }
    else
    {
      bool hit = false;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end())
        {
          pass = it->second.pass;
          fused_grid = it->second.grid;
          hit = true;
        }
      }
      if (hit)
        tuned = 2;
      else
      {
        Result r = Tune(key);
        if (r != Result::Success) return r;
        tuned = 1;
      }
    }

    // per-step tables
    std::vector<uint8_t> rec;
    std::vector<int32_t> first;
    const int grid = FusedTables(pass, fused_grid, rec, first);
    if (grid < 1) return Result::InvalidState;
    fused_grid = grid;
    Result r = Upload(fparams, rec.data(), rec.size());
    if (r == Result::Success) r = Upload(wg_first, first.data(), first.size() * sizeof(int32_t));
    if (r != Result::Success) return r;

    std::vector<ntc_dec_set> ds(n);
    std::vector<ntc_pack_set> ps(n);
    for (int s = 0; s < n; ++s)
    {
      Context::Impl& I = impl(s);
      ds[s] = {I.dw_part.as<float>(), I.dec_grad.as<float>(), I.dec_m.as<float>(), I.dec_v.as<float>(),
               I.dec.as<float>(), I.sse_part.as<double>(), I.scalars.as<double>() + 16};
      ps[s] = {I.dec.as<float>(), I.wtab.as<uint16_t>()};
    }
    std::vector<ntc_dec_step> dst((size_t)steps * n);
    const size_t tb = ntc_grid_update_tiles_bytes();
    std::vector<uint8_t> tr((size_t)steps * n * tb, 0);
    std::vector<int32_t> bf((size_t)steps * (n + 1), 0);
    // deferred sets: per (step, set) a catch-up and a gradient-step record;
    // done[s][l]: the level's steps before this one (each set's counters
    // start at zero: no set has trained yet)
    const size_t lb = ntc_grid_lazy_bytes();
    std::vector<uint8_t> cr((size_t)steps * n * lb, 0), hr((size_t)steps * n * lb, 0);
    std::vector<int32_t> cf((size_t)steps * (n + 1), 0), hf((size_t)steps * (n + 1), 0);
    std::vector<std::vector<int32_t>> done(n, std::vector<int32_t>(kMaxLevels, 0));
    lazy_cgrid = lazy_hgrid = 0;
    lazy_slots = 0;
    any_eager = false;
    step_texels.assign(steps, 0);
    const int gbudget = 8 * max_blocks;  // grid-update blocks per step, split by level size
    grid_grid = 1;
    for (int t = 0; t < steps; ++t)
    {
      std::vector<int64_t> items(n, 0);
      int64_t total = 0;
      for (int s = 0; s < n; ++s)
      {
        Context::Impl& I = impl(s);
        const StepPlan& st = I.plan[t];
        const int T = st.tile_dim;
        step_texels[t] += (int64_t)st.n_tiles * T * T;
        dst[(size_t)t * n + s] = {first[(size_t)t * (n + 1) + s + 1] - first[(size_t)t * (n + 1) + s], st.lr_decoder,
                                  0.9f, 0.999f, st.bias1, st.bias2};
        if (!st.grid_update) continue;
        const int lvl = st.level;
        const ntc_feature_desc f = I.FeatureDesc(lvl);
        const ntc_tiles tl = {I.origins_table.as<int32_t>() + (int64_t)t * 16, st.n_tiles, T, I.mip_dim[st.lod],
                              (float)st.lod / (float)(I.mip_count - 1)};
        const ntc_grid_adam adam = {I.g0m[lvl].as<float>(), I.g0v[lvl].as<float>(), I.g1m[lvl].as<float>(),
                                    I.g1v[lvl].as<float>(), st.lr_grid, 0.9f, 0.999f, st.grid_bias1,
                                    st.grid_bias2, -1.0f, 1.0f - I.g0_delta, -1.0f, 1.0f - I.g1_delta};
        if (I.lazy)
        {
          const int32_t upto = done[s][lvl]++;
          int32_t nb = 0;
          const size_t at = ((size_t)t * n + s) * lb;
          if (ntc_grid_lazy_pack(&f, &tl, nullptr, I.g0[lvl].as<float>(), I.dg0[lvl].as<float>(),
                                 I.g1[lvl].as<float>(), I.dg1[lvl].as<float>(), &adam, I.lazy_steps[lvl].as<float>(),
                                 upto, I.lazy_last[lvl][0].as<int32_t>(), I.lazy_last[lvl][1].as<int32_t>(), &hr[at],
                                 &nb) != NTC_OK)
            return Result::InvalidState;
          hf[(size_t)t * (n + 1) + s + 1] = nb;
          lazy_slots = std::max(lazy_slots, 2 * st.n_tiles);
          // the next step's catch-up, at this step's index: it runs beside
          // this step's launches. It leaves out this step's footprint cells
          // when it is on the same level (this step's fused launch reads
          // them and its gradient step brings them current) and touches
          // nothing else this step reads or writes. Step 0 is every level's
          // first visit and owes nothing; the lock step flushes every level
          // itself.
          if (t + 1 < steps && I.plan[t + 1].grid_update && !I.plan[t + 1].lock)
          {
            const StepPlan& sn = I.plan[t + 1];
            const int ln = sn.level;
            const int32_t un = done[s][ln];  // the level's steps before t + 1
            if (un > 0)
            {
              const ntc_feature_desc fn = I.FeatureDesc(ln);
              const ntc_tiles tn = {I.origins_table.as<int32_t>() + (int64_t)(t + 1) * 16, sn.n_tiles, sn.tile_dim,
                                    I.mip_dim[sn.lod], (float)sn.lod / (float)(I.mip_count - 1)};
              const ntc_grid_adam an = {I.g0m[ln].as<float>(), I.g0v[ln].as<float>(), I.g1m[ln].as<float>(),
                                        I.g1v[ln].as<float>(), sn.lr_grid, 0.9f, 0.999f, sn.grid_bias1,
                                        sn.grid_bias2, -1.0f, 1.0f - I.g0_delta, -1.0f, 1.0f - I.g1_delta};
              int32_t nc = 0;
              if (ntc_grid_lazy_pack(&fn, &tn, ln == lvl ? &tl : nullptr, I.g0[ln].as<float>(),
                                     I.dg0[ln].as<float>(), I.g1[ln].as<float>(), I.dg1[ln].as<float>(), &an,
                                     I.lazy_steps[ln].as<float>(), un, I.lazy_last[ln][0].as<int32_t>(),
                                     I.lazy_last[ln][1].as<int32_t>(), &cr[at], &nc) != NTC_OK)
                return Result::InvalidState;
              cf[(size_t)t * (n + 1) + s + 1] = nc;
              lazy_slots = std::max(lazy_slots, 2 * sn.n_tiles);
            }
          }
          continue;
        }
        any_eager = true;
        if (ntc_grid_update_tiles_pack(&f, &tl, I.g0[lvl].as<float>(), I.dg0[lvl].as<float>(), I.g1[lvl].as<float>(),
                                       I.dg1[lvl].as<float>(), &adam, &tr[((size_t)t * n + s) * tb],
                                       &items[s]) != NTC_OK)
          return Result::InvalidState;
        total += items[s];
      }
      // block offsets of the deferred sets' records (counts written above)
      for (std::vector<int32_t>* v : {&cf, &hf})
      {
        int32_t* o = &(*v)[(size_t)t * (n + 1)];
        o[0] = 0;
        for (int s = 0; s < n; ++s) o[s + 1] += o[s];
      }
      lazy_cgrid = std::max(lazy_cgrid, (int)cf[(size_t)t * (n + 1) + n]);
      lazy_hgrid = std::max(lazy_hgrid, (int)hf[(size_t)t * (n + 1) + n]);
      int32_t* f = &bf[(size_t)t * (n + 1)];
      f[0] = 0;
      for (int s = 0; s < n; ++s)
      {
        int64_t c = 0;
        if (items[s] > 0)
        {
          const int64_t need = (items[s] + 255) / 256;
          c = std::max<int64_t>(1, (int64_t)((double)gbudget * (double)items[s] / (double)total));
          c = std::min(c, need);
        }
        f[s + 1] = f[s] + (int32_t)c;
      }
      grid_grid = std::max(grid_grid, (int)f[n]);
    }
    r = Upload(dsets, ds.data(), ds.size() * sizeof(ntc_dec_set));
    if (r == Result::Success) r = Upload(psets, ps.data(), ps.size() * sizeof(ntc_pack_set));
    if (r == Result::Success) r = Upload(dsteps, dst.data(), dst.size() * sizeof(ntc_dec_step));
    if (r == Result::Success) r = Upload(trecs, tr.data(), tr.size());
    if (r == Result::Success) r = Upload(blk_first, bf.data(), bf.size() * sizeof(int32_t));
    if (r == Result::Success && lazy_hgrid > 0)
    {
      r = Upload(crecs, cr.data(), cr.size());
      if (r == Result::Success) r = Upload(cfirst, cf.data(), cf.size() * sizeof(int32_t));
      if (r == Result::Success) r = Upload(hrecs, hr.data(), hr.size());
      if (r == Result::Success) r = Upload(hfirst, hf.data(), hf.size() * sizeof(int32_t));
    }
    const int32_t zero[kSideCtr + 1] = {};
    if (r == Result::Success) r = Upload(step_ctr, zero, sizeof(zero));
    return r;
  }

  Result Train(int n_steps, bool use_graph)
  {
    // the device step counter follows step_i: the sets advance only here
    for (int s = 0; s < n; ++s)
      if (impl(s).batch_i != step_i || impl(s).phase == TrainPhase::TrainError) return Result::InvalidState;
    const int target = std::min(steps, step_i + std::max(0, n_steps));
    for (; step_i < target; ++step_i)
    {
      if (impl(0).plan[step_i].lock)
      {
        // the lock step quantises every level of every set (all sets share
        // the schedule length, so they lock together)
        if (hipStreamSynchronize(side) != hipSuccess || hipStreamSynchronize(stream) != hipSuccess)
          return Result::CudaFailure;
        for (int s = 0; s < n; ++s)
        {
          Context::Impl& I = impl(s);
          Result r = I.LockGrids();
          if (r != Result::Success) return r;
          if (hipStreamSynchronize(I.stream) != hipSuccess) return Result::CudaFailure;
        }
      }
      // the step's catch-up (issued beside the last step) before its
      // launches; this step's position for the next step's catch-up
      if (lazy_cgrid > 0 && (hipStreamWaitEvent(stream, join_ev, 0) != hipSuccess ||
                             hipEventRecord(fork_ev, stream) != hipSuccess))
        return Result::CudaFailure;
      if (use_graph)
      {
        if (!exec)
        {
          if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return Result::CudaFailure;
          const int st = LaunchStep();
          const hipError_t ce = hipStreamEndCapture(stream, &graph);
          if (st != NTC_OK || ce != hipSuccess) return Result::CudaFailure;
          if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) return Result::CudaFailure;
        }
        if (hipGraphLaunch(exec, stream) != hipSuccess) return Result::CudaFailure;
      }
      else if (LaunchStep() != NTC_OK)
        return Result::CudaFailure;
      if (lazy_cgrid > 0 && LaunchCatchup() != NTC_OK) return Result::CudaFailure;
      texels_done += step_texels[step_i];
      for (int s = 0; s < n; ++s)
      {
        Context::Impl& I = impl(s);
        const StepPlan& st = I.plan[step_i];
        I.grid_batch_i[st.level] += 1;
        // the deferred sets' step counters, as Context::Train keeps them (the
        // lock and the evaluation flush from them)
        if (I.lazy && st.grid_update)
        {
          I.lazy_done[st.level] += 1;
          I.lazy_behind[st.level] = true;
        }
      }
    }
    if (hipStreamSynchronize(side) != hipSuccess || hipStreamSynchronize(stream) != hipSuccess)
      return Result::CudaFailure;
    for (int s = 0; s < n; ++s)
    {
      Context::Impl& I = impl(s);
      I.batch_i = step_i;
      ++I.state_version;
      if (step_i == steps && I.phase != TrainPhase::TrainComplete)
      {
        I.phase = TrainPhase::TrainComplete;
        Result r = I.Finalize(sets[s]->ctx);
        if (r != Result::Success) return r;
      }
    }
    return Result::Success;
  }
};

extern "C" {
int ntc_batch_create(ntc_context* const* sets, int32_t n_sets, int32_t pass_texels, int32_t fused_grid,
                     ntc_batch** out)
{
  if (!out || !sets || n_sets < 1 || n_sets > 64) return (int)Result::InvalidState;
  *out = nullptr;
  ntc_batch* b = new ntc_batch();
  for (int s = 0; s < n_sets; ++s)
  {
    if (!sets[s])
    {
      delete b;
      return (int)Result::InvalidState;
    }
    for (int q = 0; q < s; ++q)
      if (sets[q] == sets[s])
      {
        delete b;
        return (int)Result::InvalidState;
      }
    b->sets.push_back(sets[s]);
  }
  b->n = n_sets;
  const Result r = b->Setup(pass_texels, fused_grid);
  if (r != Result::Success)
  {
    delete b;
    return (int)r;
  }
  *out = b;
  return 0;
}

int ntc_batch_train(ntc_batch* b, int32_t n_steps, int32_t use_graph)
{
  if (!b) return (int)Result::InvalidState;
  return (int)b->Train(n_steps, use_graph != 0);
}

int ntc_batch_get_stats(const ntc_batch* b, ntc_batch_stats* out)
{
  if (!b || !out) return (int)Result::InvalidState;
  std::memset(out, 0, sizeof(*out));
  out->n_sets = b->n;
  out->steps = b->steps;
  out->steps_done = b->step_i;
  out->pass_texels = b->pass;
  out->fused_grid = b->fused_grid;
  out->grid_update_grid = b->grid_grid;
  out->tuned = b->tuned;
  out->texels_done = b->texels_done;
  for (int i = 0; i < 4; ++i)
  {
    out->tune_ms[i] = b->tune_ms[i];
    out->tune_pass[i] = b->tune_pass[i];
    out->tune_grid[i] = b->tune_grid[i];
  }
  out->graph = b->exec != nullptr;
  return 0;
}

void ntc_batch_destroy(ntc_batch* b) { delete b; }
} // extern "C"
