// rtm_driver.cpp -- the reference's main() flow (kernel.cu:525-1213) on top of
// librtm_hip.so, with a pthread shot dispatcher.
//
// Flow: read the two parameter files and the source depths; pad the model;
// snap the velocity range and design the coefficient table; then shots are
// grouped into device batches and handed to worker threads.  Each worker owns
// one librtm_hip context (its own HIP stream and device buffers) and runs
// forward pass, backward pass with imaging and the per-shot image post for
// its batch, writing RVSP_RTM_up_<m>.dat / RVSP_RTM_down_<m>.dat exactly as
// kernel.cu:951-990 does.  Those files are the checkpoints: a rerun skips
// shots whose two files exist.  A failed batch is re-queued up to `retries`
// times.  Finally the per-shot images are stacked in shot order
// (kernel.cu:1000-1059), cropped (kernel.cu:1061-1108), taken through the
// depth-to-time / phase-rotation / time-to-depth chain into its three side
// files (kernel.cu:1110-1179) and written as SEG-Y (kernel.cu:1181-1209).
#include <pthread.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/rtm_hip.h"
#include "../../../include/rtm_host.h"
#include "rtm_host_internal.h"

namespace {

using rtmh::set_error;

struct Batch {
  std::vector<int> shots;  // 0-based source indices m
  int attempts = 0;
};

struct Job {
  rtmh_config cfg;
  rtmh_derived d;
  rtmh_run_opts opt;
  std::vector<float> v, r1, w, c, wav, depths;
  std::vector<int> index;
  int nvel = 0;
  float vmin = 0, vmax = 0;
  std::mutex mu;
  std::deque<Batch> queue;
  int failures = 0;
  std::string first_error;
  int done_shots = 0;
};

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

std::string shot_file(const Job& J, const char* kind, int m) {
  char name[64];
  snprintf(name, sizeof(name), "RVSP_RTM_%s_%d.dat", kind, m + 1);
  return std::string(J.cfg.result) + name;
}

// Gather file of kernel.cu:827-838: <OutNameseis>NEW_L10-1932-X_<N>.dat with
// N the integer depth; n traces of NT1 floats.
int read_gather(const Job& J, int m, float* out) {
  char name[64];
  snprintf(name, sizeof(name), "NEW_L10-1932-X_%d.dat", (int)J.depths[m]);
  std::string p = std::string(J.cfg.out_seis) + name;
  FILE* f = fopen(p.c_str(), "rb");
  if (!f) return set_error(rtmh::RTMH_E_IO, "cannot open gather %s", p.c_str());
  const size_t n = (size_t)J.cfg.n * J.cfg.nt1;
  const size_t got = fread(out, sizeof(float), n, f);
  fclose(f);
  if (got != n) return set_error(rtmh::RTMH_E_FORMAT, "%s: %zu of %zu floats", p.c_str(), got, n);
  return rtmh::RTMH_OK;
}

int write_interior_colmajor(const std::string& path, const float* img, int mnz, int mnx) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return set_error(rtmh::RTMH_E_IO, "cannot create %s", path.c_str());
  std::vector<float> col(mnz);
  for (int x = 0; x < mnx; ++x) {  // j outer, i inner (kernel.cu:954-959)
    for (int z = 0; z < mnz; ++z) col[z] = img[(size_t)z * mnx + x];
    fwrite(col.data(), sizeof(float), mnz, f);
  }
  fclose(f);
  return rtmh::RTMH_OK;
}

int run_batch(Job& J, rtm_ctx* ctx, const Batch& b) {
  const rtmh_config& c = J.cfg;
  const int nb = (int)b.shots.size();
  const int mnz = c.mod_nz, mnx = c.mod_nx;
  std::vector<int> ru(nb), rx(nb, J.d.r_x);
  std::vector<float> seis((size_t)nb * c.n * J.d.nt);
  for (int i = 0; i < nb; ++i) {
    const int m = b.shots[i];
    ru[i] = rtmh_source_row(J.depths[m], c.hz, c.n2);
    int rc = read_gather(J, m, &seis[(size_t)i * c.n * J.d.nt]);
    if (rc) return rc;
  }
  int rc = rtm_set_shots(ctx, nb, ru.data(), rx.data());
  if (!rc) rc = rtm_set_gathers(ctx, seis.data());
  if (!rc) rc = rtm_migrate(ctx);
  std::vector<float> up((size_t)nb * mnz * mnx), down((size_t)nb * mnz * mnx);
  if (!rc) rc = rtm_image_post(ctx, J.vmax, c.whitecoe, up.data(), down.data());
  if (rc) return set_error(rtmh::RTMH_E_DEVICE, "device batch failed (%d): %s", rc, rtm_last_error(ctx));
  for (int i = 0; i < nb; ++i) {
    const int m = b.shots[i];
    rc = write_interior_colmajor(shot_file(J, "up", m), &up[(size_t)i * mnz * mnx], mnz, mnx);
    if (!rc) rc = write_interior_colmajor(shot_file(J, "down", m), &down[(size_t)i * mnz * mnx], mnz, mnx);
    if (rc) return rc;
  }
  return rtmh::RTMH_OK;
}

void* worker(void* arg) {
  Job& J = *static_cast<Job*>(arg);
  const rtmh_config& c = J.cfg;
  rtm_params p{};
  p.nz = J.d.nz; p.nx = J.d.nx; p.n2 = c.n2; p.nfdmax = c.nfdmax; p.nt = J.d.nt; p.nt2 = J.d.nt2;
  p.ilste = c.ilste; p.icompen = c.icompen;
  p.s_l = J.d.s_l; p.s_r = J.d.s_r; p.s_z = J.d.s_z; p.ds = c.ds; p.n = c.n;
  p.tao2 = J.d.tao2; p.h2 = J.d.h2; p.taoh = J.d.taoh; p.taoh2 = J.d.taoh2; p.hzx2_1 = J.d.hzx2_1;
  p.max_shots = J.opt.batch; p.strip_format = J.opt.strip_format; p.record = 0; p.zseg = 0;
  rtm_ctx* ctx = nullptr;
  int rc = rtm_create(&p, &ctx);
  if (!rc)
    rc = rtm_set_model(ctx, J.v.data(), J.r1.data(), J.w.data(), J.c.data(), (int)J.c.size(),
                       c.ilste == 0 ? J.index.data() : nullptr, J.nvel, J.vmin, c.dv);
  if (!rc) rc = rtm_set_wavelet(ctx, J.wav.data());
  if (rc) {
    std::lock_guard<std::mutex> g(J.mu);
    J.failures++;
    if (J.first_error.empty())
      J.first_error = std::string("context setup failed: ") + (ctx ? rtm_last_error(ctx) : "rtm_create");
    if (ctx) rtm_destroy(ctx);
    return nullptr;
  }
  for (;;) {
    Batch b;
    {
      std::lock_guard<std::mutex> g(J.mu);
      if (J.queue.empty()) break;
      b = J.queue.front();
      J.queue.pop_front();
    }
    rc = 0;
    if (J.opt.before_batch) {
      const int hr = J.opt.before_batch(b.shots.data(), (int)b.shots.size(), b.attempts, J.opt.user);
      if (hr) rc = set_error(rtmh::RTMH_E_REFUSED, "attempt %d of the batch from shot %d refused (%d)",
                             b.attempts, b.shots[0] + 1, hr);
    }
    if (!rc) rc = run_batch(J, ctx, b);
    std::lock_guard<std::mutex> g(J.mu);
    if (rc) {
      if (++b.attempts <= J.opt.retries) {
        J.queue.push_back(b);
      } else {
        J.failures++;
        if (J.first_error.empty()) J.first_error = rtmh_last_error();
      }
    } else {
      J.done_shots += (int)b.shots.size();
      if (J.opt.verbose) fprintf(stderr, "rtm_driver: %d shots done\n", J.done_shots);
    }
  }
  rtm_destroy(ctx);
  return nullptr;
}

}  // namespace

extern "C" int rtmh_run(const char* rtm_txt, const rtmh_run_opts* opts) {
  Job J;
  J.opt = opts ? *opts : rtmh_run_opts{1, 8, 1, 0, 1, 0, nullptr, nullptr, nullptr};
  if (J.opt.workers < 1) J.opt.workers = 1;
  if (J.opt.batch < 1) J.opt.batch = 1;
  int rc = rtmh_read_config(rtm_txt, &J.cfg);
  if (rc) return rc;
  rtmh_config& c = J.cfg;
  rtmh_derive(&c, &J.d);
  if (J.d.nt != c.nt1)
    return set_error(rtmh::RTMH_E_ARG, "NT=%d differs from NT1=%d: trace resampling is not supported", J.d.nt, c.nt1);
  J.depths.resize(c.nrec);
  if ((rc = rtmh_read_depths(c.out_dpr, c.nrec, J.depths.data()))) return rc;
  std::vector<float> raw((size_t)c.mod_nz * c.mod_nx);
  if ((rc = rtmh_read_model(c.out_vp, c.mod_nz, c.mod_nx, raw.data()))) return rc;
  const size_t nxz = (size_t)J.d.nz * J.d.nx;
  J.v.resize(nxz);
  J.r1.resize(nxz);
  rtmh_pad_model(raw.data(), c.mod_nz, c.mod_nx, c.n2, c.ifv, c.tao, c.h, J.v.data(), J.r1.data());
  rtmh_vrange(J.v.data(), (long long)nxz, c.dv, &J.vmin, &J.vmax, &J.nvel);
  J.w.resize(c.n2 + 1);
  rtmh_blend_weights(c.n2, J.w.data());
  J.wav.resize(J.d.nt + 1);
  rtmh_wavelet_table(J.d.nt, c.tao, c.f0, J.wav.data());
  if (c.ilste == 0) {
    std::vector<int> need(J.nvel), M(J.nvel);
    J.index.resize(J.nvel + 1);
    rtmh_bins_needed(J.v.data(), (long long)nxz, J.vmin, c.dv, J.nvel, need.data());
    J.c.resize((size_t)J.nvel * (c.nfdmax + 1));
    int nc = rtmh_design_ls(c.nthita, c.nfdmax, c.nfdmin, J.nvel, c.tao, c.h, c.df, c.eps, c.fmax,
                            J.vmin, c.dv, need.data(), c.hz / c.h, M.data(), J.index.data(),
                            J.c.data(), (int)J.c.size(), 16);
    if (nc < 0) return nc;
    J.c.resize(nc);
  } else {
    J.c.resize(c.nfdmax + 1);
    rtmh_taylor(c.nfdmax, J.c.data());
  }
  {  // the reference runs any time step; an unstable one only yields NaN images
    const double st = rtmh_stability(J.v.data(), (long long)nxz, J.c.data(),
                                     c.ilste == 0 ? J.index.data() : nullptr, J.nvel, J.vmin, c.dv,
                                     c.nfdmax, J.d.tao2, J.d.h2, J.d.hzx2_1);
    if (st > 1.0)
      return set_error(rtmh::RTMH_E_UNSTABLE,
                       "time step %g s exceeds the stability limit of the model (number %.4f > 1; "
                       "the largest stable step is about %g s)", (double)c.tao, st,
                       (double)c.tao / std::sqrt(st));
  }
  // batches of shots still to do
  Batch cur;
  for (int m = 0; m < c.nrec; ++m) {
    if (J.opt.resume && file_exists(shot_file(J, "up", m)) && file_exists(shot_file(J, "down", m))) continue;
    cur.shots.push_back(m);
    if ((int)cur.shots.size() == J.opt.batch) {
      J.queue.push_back(cur);
      cur.shots.clear();
    }
  }
  if (!cur.shots.empty()) J.queue.push_back(cur);
  const int nw = std::min<int>(J.opt.workers, std::max<size_t>(J.queue.size(), 1));
  std::vector<pthread_t> th(nw);
  for (int i = 0; i < nw; ++i) pthread_create(&th[i], nullptr, worker, &J);
  for (int i = 0; i < nw; ++i) pthread_join(th[i], nullptr);
  if (J.failures) return set_error(rtmh::RTMH_E_DEVICE, "%d batch(es) failed: %s", J.failures, J.first_error.c_str());

  // stack the per-shot files in shot order (kernel.cu:1000-1059)
  const int mnz = c.mod_nz, mnx = c.mod_nx;
  const size_t nI = (size_t)mnz * mnx;
  std::vector<float> s1(nI, 0.f), s2(nI, 0.f), col(nI);
  for (int m = 0; m < c.nrec; ++m) {
    for (int k = 0; k < 2; ++k) {
      std::string p = shot_file(J, k ? "down" : "up", m);
      FILE* f = fopen(p.c_str(), "rb");
      if (!f || fread(col.data(), sizeof(float), nI, f) != nI) {
        if (f) fclose(f);
        return set_error(rtmh::RTMH_E_IO, "cannot read %s", p.c_str());
      }
      fclose(f);
      std::vector<float>& s = k ? s2 : s1;
      for (int x = 0; x < mnx; ++x)
        for (int z = 0; z < mnz; ++z) s[(size_t)z * mnx + x] += col[(size_t)x * mnz + z];
    }
  }
  std::vector<float> img(nI);
  for (size_t i = 0; i < nI; ++i) {
    float a = s1[i] / c.nrec, b = s2[i] / c.nrec;
    img[i] = (c.inorm == 1) ? a / b : a;
  }
  // crop and flip (kernel.cu:1061-1108); the image is in interior coordinates
  const std::string res(c.result);
  std::vector<float> crop;
  FILE* f = fopen((res + "RVSP_Migration_Real_new2.dat").c_str(), "wb");
  if (!f) return set_error(rtmh::RTMH_E_IO, "cannot create output in %s", c.result);
  auto put_img = [&](int z, int x) {
    float val = (z >= 0 && z < mnz && x >= 0 && x < mnx) ? img[(size_t)z * mnx + x] : 0.f;
    crop.push_back(val);
  };
  if (c.ifv == 1) {
    for (int j = c.nx_ed; j > c.nx_bg; --j)
      for (int i = c.nz_bg; i < c.nz_ed; ++i) put_img(i, j);
  } else {
    for (int j = c.nx_bg; j < c.nx_ed; ++j)
      for (int i = c.nz_bg; i < c.nz_ed; ++i) put_img(i, j);
  }
  fwrite(crop.data(), sizeof(float), crop.size(), f);
  fclose(f);
  // the velocity window beside it (vnew.dat, kernel.cu:1074-1107), padded
  // coordinates; note the ifv = 0 loop starts at NZ_BG, not NZ_BG + N2
  std::vector<float> vw;
  {
    auto put_v = [&](int z, int x) {
      vw.push_back((z >= 0 && z < J.d.nz && x >= 0 && x < J.d.nx) ? J.v[(size_t)z * J.d.nx + x] : 0.f);
    };
    if (c.ifv == 1) {
      for (int j = c.nx_ed + c.n2; j > c.nx_bg + c.n2; --j)
        for (int i = c.nz_bg + c.n2; i < c.nz_ed + c.n2; ++i) put_v(i, j);
    } else {
      for (int j = c.nx_bg + c.n2; j < c.nx_ed + c.n2; ++j)
        for (int i = c.nz_bg; i < c.nz_ed + c.n2; ++i) put_v(i, j);
    }
    FILE* fv = fopen((res + "vnew.dat").c_str(), "wb");
    if (!fv) return set_error(rtmh::RTMH_E_IO, "cannot create vnew.dat in %s", c.result);
    fwrite(vw.data(), sizeof(float), vw.size(), fv);
    fclose(fv);
  }
  // post-stack chain (kernel.cu:1110-1179): both files are read back as
  // SGY_N columns of mod_NZ samples, in file order; the image goes to two-way
  // time, is rotated by `angle` degrees and goes back to depth.
  const int sgy_n = c.nx_ed - c.nx_bg;
  if (sgy_n > 0) {
    const size_t need = (size_t)sgy_n * mnz;
    std::vector<float> mig(need, 0.f), vdt(need, 0.f);
    std::copy_n(crop.begin(), std::min(crop.size(), need), mig.begin());
    std::copy_n(vw.begin(), std::min(vw.size(), need), vdt.begin());
    auto write_file = [&](const char* name, const std::vector<float>& a) {
      FILE* fo = fopen((res + name).c_str(), "wb");
      if (!fo) return set_error(rtmh::RTMH_E_IO, "cannot create %s in %s", name, c.result);
      fwrite(a.data(), sizeof(float), a.size(), fo);
      fclose(fo);
      return (int)rtmh::RTMH_OK;
    };
    const int ntime = rtmh_d2t(vdt.data(), mig.data(), sgy_n, mnz, 0, sgy_n - 1, c.hz, c.tao, nullptr, 0);
    if (ntime < 0) return ntime;
    std::vector<float> mt((size_t)sgy_n * ntime), mtp((size_t)sgy_n * ntime);
    rc = rtmh_d2t(vdt.data(), mig.data(), sgy_n, mnz, 0, sgy_n - 1, c.hz, c.tao, mt.data(), (long long)mt.size());
    if (rc < 0) return rc;
    if ((rc = write_file("RVSP_Migration_Real_T.dat", mt)) != 0) return rc;
    rtmh_phase_rotate(mt.data(), mtp.data(), sgy_n, ntime, c.angle);
    if ((rc = write_file("RVSP_Migration_Real_T_phase.dat", mtp)) != 0) return rc;
    const int ndep = rtmh_t2d(vdt.data(), mtp.data(), sgy_n, ntime, mnz, 0, sgy_n - 1, c.tao, c.hz, nullptr, 0);
    if (ndep < 0) return ndep;
    std::vector<float> md((size_t)sgy_n * ndep);
    rc = rtmh_t2d(vdt.data(), mtp.data(), sgy_n, ntime, mnz, 0, sgy_n - 1, c.tao, c.hz, md.data(), (long long)md.size());
    if (rc < 0) return rc;
    if ((rc = write_file("RVSP_Migration_Real_D.dat", md)) != 0) return rc;
  }
  // SEG-Y output (kernel.cu:1181-1209): SGY_N traces of mod_NZ samples read
  // sequentially from the cropped image.
  if (sgy_n > 0) {
    std::vector<float> data((size_t)sgy_n * mnz, 0.f), sx(sgy_n), sy(sgy_n), dsr(sgy_n);
    std::copy_n(crop.begin(), std::min(crop.size(), data.size()), data.begin());
    for (int i = 0; i < sgy_n; ++i) {
      sx[i] = (float)(1.0 * i * 10);
      sy[i] = (float)(-1.0 * i * 10);
      dsr[i] = (float)(1000 - i);
    }
    const char* tpl = J.opt.sgy_template ? J.opt.sgy_template : "SGY_Model.sgy";
    rc = rtmh_write_segy(tpl, (res + "RVSP_migration_Real.sgy").c_str(), data.data(), sgy_n, mnz,
                         (int)c.hz, sx.data(), sy.data(), 1.0f, 1.0f, dsr.data());
    if (rc) return rc;
  }
  return rtmh::RTMH_OK;
}
