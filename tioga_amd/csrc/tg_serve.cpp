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
// A request's moved coordinates are computed on the device from the moving
// block's resident coordinates (and on the host for its wall nodes only,
// which the hole map's wall bounds read), and the block is preprocessed
// there (device OBB with the bits of the host running sums).
//
// Resident-tree mode: the background's search structure is one tree over
// all its cells, built once when serving opens, instead of the tree over
// the cells overlapping each request's query set.  Every cell whose padded
// box holds a query point overlaps that query set's OBB, so both trees offer
// a point the same candidate cells, in different walk orders; the donor
// differs only where it depends on the order (more than one candidate
// accepts the point, or one accepts it with a near-zero weight in a BIGVALUE
// cell).  Those queries are counted per request (tieband); with a count of
// 0 the request's results are those of the per-request tree.
//
// Admission: every device allocation a slot's engine makes is counted on
// the slot's thread (DBuf and the primitives' scratch).  A slot holds
// `retained` bytes between requests; a request is charged the most any
// request on its moving block has held so far (its first one: bytes_per_cell
// per cell plus bytes_per_node per node), or the slot's retained bytes if
// more.  A request starts, in arrival order, when its charge plus every
// other slot's (running: its charge, idle: its retained bytes) fits the
// budget; one request always runs.  Each result reports its charge and the
// bytes its slot actually held at most.  A slot's buffers only grow, so a
// request can hold a little more than its charge where the slot kept some
// buffers larger from requests on other blocks; the charge is a measured
// estimate, not a bound.
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

Server::Server(Engine &main, int bg_tag, int nslots, double budget_bytes, bool resident_tree)
    : main_(main), bg_tag_(bg_tag), budget_(budget_bytes), resident_(resident_tree)
{
  if (nslots < 1 || nslots > 64) throw std::runtime_error("tioga_amd: serving needs 1..64 slots");
  if (main.block_nnodes(bg_tag) < 0) throw std::runtime_error("tioga_amd: unknown background tag");
  if (resident_) main.build_resident_tree(bg_tag, tree_);
  for (int i = 0; i < nslots; i++) {
    Slot *s = new Slot;
    s->eng.isym = main.isym;
    s->eng.nfringe = main.nfringe;
    s->eng.mexclude = main.mexclude;
    s->eng.batched = true;
    s->eng.overlap_dedupe = false;
    s->eng.adopt_block(main, bg_tag, bg_tag, 2, nullptr, nullptr, false);
    if (resident_) s->eng.use_resident_tree(bg_tag, &tree_);
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

double Server::estimate(int tag) const
{
  auto it = seen_peak_.find(tag);
  const double formula = bytes_per_cell * (double)main_.block_ncells(tag) + bytes_per_node * (double)main_.block_nnodes(tag);
  return it == seen_peak_.end() ? formula : std::max(formula, it->second);
}

void Server::worker(Slot &s)
{
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      auto fits = [&] {
        if (queue_.empty()) return false;
        if (active_ == 0) return true;
        double others = 0;
        for (Slot *o : slots_)
          if (o != &s) others += o->charge;
        return others + std::max(s.retained, estimate(queue_.front()->tag)) <= budget_;
      };
      cv_work_.wait(lk, [&] { return stop_ || fits(); });
      if (queue_.empty()) return;  // stop_ with nothing left
      r = queue_.front();
      queue_.pop_front();
      r->est = std::max(s.retained, estimate(r->tag));
      s.charge = r->est;
      active_++;
    }
    DevBytes &dev = thread_dev_bytes();
    dev.peak = dev.now;
    r->res.t_start = now_s();
    try {
      run(s, *r);
      r->res.status = 0;
    } catch (const std::exception &e) {
      r->res.status = 1;
      r->res.err = e.what();
    }
    r->res.t_done = now_s();
    r->res.est_bytes = r->est;
    r->res.peak_bytes = (double)dev.peak;
    {
      std::lock_guard<std::mutex> lk(mu_);
      double &sp = seen_peak_[r->tag];
      sp = std::max(sp, (double)dev.peak);
      s.retained = (double)dev.now;
      s.charge = s.retained;
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
  // x' = R (x - c) + c + t, per node and in this order: the wall nodes on the
  // host (the hole map's wall bounds read them), all nodes on the device
  int nwbc = 0;
  const int *wbc = main_.block_wall_nodes(r.tag, &nwbc);
  for (int k = 0; k < nwbc; k++) {
    const size_t i = (size_t)(wbc[k] - BASE);
    const double d0 = x0[3 * i] - r.c[0], d1 = x0[3 * i + 1] - r.c[1], d2 = x0[3 * i + 2] - r.c[2];
    for (int j = 0; j < 3; j++) {
      const double y = r.rot[3 * j] * d0 + r.rot[3 * j + 1] * d1 + r.rot[3 * j + 2] * d2;
      s.x[3 * i + j] = y + r.c[j] + r.t[j];
    }
  }
  s.eng.adopt_block(main_, r.tag, r.tag, 1, s.x.data(), s.ib.data(), true, s.eng.block_count() > 1 ? 1 : -1);
  s.eng.profile_moved(r.tag, main_.block_device_coords(r.tag), r.rot, r.c, r.t);
  s.eng.perform_connectivity();
  r.res.tieband = s.eng.last_tieband();
  r.res.shell_iblank.assign(s.ib.begin(), s.ib.begin() + n);
  s.eng.nonunit_iblank(bg_tag_, r.res.bg_nodes, r.res.bg_vals);
  r.res.nsearch = s.eng.last_times.nsearch;
  r.res.connectivity_ms = s.eng.last_times.total;
}

}  // namespace tg
