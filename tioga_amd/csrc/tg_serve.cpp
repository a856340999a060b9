// tg_serve.cpp -- a request loop for moving bodies over a resident background.
//
// A request moves one of the registered near-body blocks rigidly (rotation
// R about a centre c, then a shift t: x' = R (x - c) + c + t, evaluated per
// node in that order) and asks for the connectivity of that moved block with
// the background block alone: exactly what tioga::performConnectivity
// (tioga.C:112-160) returns for that two-block system, run in the batched
// form.
//
// Requests are served by slots.  A slot is an engine on its own stream whose
// background block is a view of the main engine's (coordinates, topology,
// resolutions, bins and OBB stay resident and are never copied) and whose
// moving block shares the main engine's topology; everything a connectivity
// writes (iblank, search arrays, ADTs, hole map, donor tables) is the slot's
// own.  Slots run concurrently on host threads, so requests that arrive while
// others are on the GPU start as soon as a slot is free and the HBM budget
// admits them, and their work overlaps on the device.  A slot's result
// depends on its request alone.
//
// Admission: a request's device memory beyond the slot's fixed part is
// estimated from the moving block's size (bytes_per_cell per cell plus
// bytes_per_node per node); requests start in arrival order while the
// estimates of the running ones plus the next one fit the budget (one
// request always runs, whatever its estimate).
#include "tg_serve.h"

#include <chrono>
#include <cstring>
#include <stdexcept>

namespace tg {

namespace {

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

Server::Server(Engine &main, int bg_tag, int nslots, double budget_bytes)
    : main_(main), bg_tag_(bg_tag), budget_(budget_bytes)
{
  if (nslots < 1 || nslots > 64) throw std::runtime_error("tioga_amd: serving needs 1..64 slots");
  if (main.block_nnodes(bg_tag) < 0) throw std::runtime_error("tioga_amd: unknown background tag");
  for (int i = 0; i < nslots; i++) {
    Slot *s = new Slot;
    s->eng.isym = main.isym;
    s->eng.nfringe = main.nfringe;
    s->eng.mexclude = main.mexclude;
    s->eng.batched = true;
    s->eng.overlap_dedupe = false;
    s->eng.adopt_block(main, bg_tag, bg_tag, 2, nullptr, nullptr, false);
    slots_.push_back(s);
  }
  for (Slot *s : slots_) s->th = std::thread([this, s] { worker(*s); });
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (Slot *s : slots_) {
    if (s->th.joinable()) s->th.join();
    delete s;
  }
  for (auto &kv : reqs_) delete kv.second;
}

long Server::submit(int shell_tag, const double rot[9], const double center[3], const double shift[3])
{
  const int n = main_.block_nnodes(shell_tag);
  if (n < 0 || shell_tag == bg_tag_) throw std::runtime_error("tioga_amd: request for an unknown moving block");
  Req *r = new Req;
  r->tag = shell_tag;
  memcpy(r->rot, rot, sizeof r->rot);
  memcpy(r->c, center, sizeof r->c);
  memcpy(r->t, shift, sizeof r->t);
  r->est = bytes_per_cell * (double)main_.block_ncells(shell_tag) + bytes_per_node * (double)n;
  r->res.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(mu_);
    r->id = next_id_++;
    reqs_[r->id] = r;
    queue_.push_back(r);
  }
  cv_work_.notify_all();
  return r->id;
}

void Server::wait(long id, ServeResult &out)
{
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::runtime_error("tioga_amd: unknown request id");
  Req *r = it->second;
  cv_done_.wait(lk, [&] { return r->done; });
  out = std::move(r->res);
  reqs_.erase(it);
  delete r;
}

void Server::worker(Slot &s)
{
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] {
        return stop_ || (!queue_.empty() && (active_ == 0 || reserved_ + queue_.front()->est <= budget_));
      });
      if (queue_.empty()) return;  // stop_ with nothing left
      r = queue_.front();
      queue_.pop_front();
      reserved_ += r->est;
      active_++;
    }
    r->res.t_start = now_s();
    try {
      run(s, *r);
      r->res.status = 0;
    } catch (const std::exception &e) {
      r->res.status = 1;
      r->res.err = e.what();
    }
    r->res.t_done = now_s();
    {
      std::lock_guard<std::mutex> lk(mu_);
      reserved_ -= r->est;
      active_--;
      r->done = true;
    }
    cv_done_.notify_all();
    cv_work_.notify_all();
  }
}

void Server::run(Slot &s, Req &r)
{
  const int n = main_.block_nnodes(r.tag);
  const double *x0 = main_.block_coords(r.tag);
  s.x.resize(3 * (size_t)n);
  s.ib.resize(std::max(n, 1));
  // x' = R (x - c) + c + t, per node and in this order
  for (int i = 0; i < n; i++) {
    const double d0 = x0[3 * (size_t)i] - r.c[0], d1 = x0[3 * (size_t)i + 1] - r.c[1],
                 d2 = x0[3 * (size_t)i + 2] - r.c[2];
    for (int j = 0; j < 3; j++) {
      const double y = r.rot[3 * j] * d0 + r.rot[3 * j + 1] * d1 + r.rot[3 * j + 2] * d2;
      s.x[3 * (size_t)i + j] = y + r.c[j] + r.t[j];
    }
  }
  s.eng.adopt_block(main_, r.tag, r.tag, 1, s.x.data(), s.ib.data(), true, s.eng.block_count() > 1 ? 1 : -1);
  s.eng.profile();
  s.eng.perform_connectivity();
  r.res.shell_iblank.assign(s.ib.begin(), s.ib.begin() + n);
  s.eng.nonunit_iblank(bg_tag_, r.res.bg_nodes, r.res.bg_vals);
  r.res.nsearch = s.eng.last_times.nsearch;
  r.res.connectivity_ms = s.eng.last_times.total;
}

}  // namespace tg
