// tg_serve.h -- request loop for moving bodies over a resident background
// (tg_serve.cpp).
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tg_engine.h"

namespace tg {

struct ServeResult {
  int status = 0;  // 0 done, 1 failed (err says why)
  std::string err;
  std::vector<int> shell_iblank;     // iblank of every node of the moved block
  std::vector<int> bg_nodes, bg_vals;  // background nodes whose iblank is not 1
  long long nsearch = 0;
  long long tieband = 0;     // resident-tree mode: queries whose donor could depend on the walk order
  double est_bytes = 0;      // device memory the admission charged for the request
  double peak_bytes = 0;     // device memory the slot held at most during the request
  double connectivity_ms = 0;
  double t_submit = 0, t_start = 0, t_done = 0;  // seconds, steady clock
};

class Server {
 public:
  // main: an engine whose blocks are registered and preprocessed; bg_tag the
  // resident background; nslots concurrent requests at most; budget_bytes
  // the device memory the slots may hold together (see Admission in
  // tg_serve.cpp); resident_tree: searches against the background walk one
  // tree over all its cells, built here, instead of a tree per request.
  Server(Engine &main, int bg_tag, int nslots, double budget_bytes, bool resident_tree = false);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  long submit(int shell_tag, const double rot[9], const double center[3], const double shift[3]);
  void wait(long id, ServeResult &out);

  double bytes_per_cell = 1536.0, bytes_per_node = 512.0;

 private:
  struct Req {
    long id = 0;
    int tag = 0;
    double rot[9], c[3], t[3];
    double est = 0;
    ServeResult res;
    bool done = false;
  };
  struct Slot {
    Engine eng;
    std::vector<double> x;
    std::vector<int> ib;
    std::thread th;
    double retained = 0;  // device bytes the slot holds between requests
    double charge = 0;    // what the admission counts for the slot now
  };
  void worker(Slot &s);
  void run(Slot &s, Req &r);
  double estimate(int tag) const;  // under mu_

  Engine &main_;
  int bg_tag_;
  double budget_;
  bool resident_;
  ResidentTree tree_;
  std::map<int, double> seen_peak_;  // per moving block: the most a request has held
  int active_ = 0;
  long next_id_ = 1;
  bool stop_ = false;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Req *> queue_;
  std::map<long, Req *> reqs_;
  std::vector<Slot *> slots_;
};

}  // namespace tg
