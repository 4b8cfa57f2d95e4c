// Single-process stand-in for the MPI C++ bindings the reference's drivers
// and headers use (MPI::Init, MPI::COMM_WORLD, Comm::Get_rank, Reduce, ...).
//
// The MI355X path runs one process per GPU and this build drives one GPU,
// so every communicator has one rank: collectives copy the send buffer to
// the receive buffer (or leave MPI::IN_PLACE data alone), Barrier returns,
// and point-to-point calls, which need a second rank, exit with a message.
// Drivers compiled against include/bcm see this header instead of an MPI
// installation's; nothing is linked for it.
#ifndef BCM_FACADE_MPI_H
#define BCM_FACADE_MPI_H

#include <sys/time.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace MPI {

// A datatype is its element size.
class Datatype {
  int bytes;

 public:
  explicit Datatype(int b = 0) : bytes(b) {}
  int Get_size() const { return bytes; }
};

class Op {
  int code;

 public:
  explicit Op(int c = 0) : code(c) {}
  int Get_code() const { return code; }
};

static const Datatype CHAR(1), UNSIGNED_CHAR(1), BYTE(1), SHORT(2), INT(4), UNSIGNED(4), LONG(8),
    UNSIGNED_LONG(8), LONG_LONG(8), FLOAT(4), DOUBLE(8), LONG_DOUBLE(sizeof(long double)), BOOL(sizeof(bool));
static const Op SUM(1), PROD(2), MAX(3), MIN(4), LAND(5), LOR(6), BAND(7), BOR(8), MAXLOC(9), MINLOC(10);

static const int PROC_NULL = -2;
static const int ANY_SOURCE = -1;
static const int ANY_TAG = -1;
static const int UNDEFINED = -32766;

// MPI::IN_PLACE: an address no buffer can have, the same in every
// translation unit.
static void* const IN_PLACE = reinterpret_cast<void*>(~std::uintptr_t(0));

inline void single_rank_only(const char* what) {
  std::printf("*** error: %s needs a second MPI rank; this build runs one process on one GPU\n", what);
  std::exit(1);
}

class Status {
 public:
  int Get_source() const { return 0; }
  int Get_tag() const { return 0; }
  int Get_count(const Datatype&) const { return 0; }
};

class Request {
 public:
  void Wait() {}
  void Wait(Status&) {}
  bool Test() { return true; }
  static void Waitall(int, Request*) {}
  static void Waitall(int, Request*, Status*) {}
};

class Comm {
 protected:
  static void copy(const void* send, void* recv, int count, const Datatype& t) {
    if (send != IN_PLACE && send != recv && count > 0)
      std::memmove(recv, send, size_t(count) * size_t(t.Get_size()));
  }

 public:
  virtual ~Comm() {}
  int Get_rank() const { return 0; }
  int Get_size() const { return 1; }
  void Barrier() const {}
  void Abort(int code) const { std::exit(code); }
  void Bcast(void*, int, const Datatype&, int) const {}
  void Reduce(const void* send, void* recv, int count, const Datatype& t, const Op&, int) const {
    copy(send, recv, count, t);
  }
  void Allreduce(const void* send, void* recv, int count, const Datatype& t, const Op&) const {
    copy(send, recv, count, t);
  }
  void Gather(const void* send, int scount, const Datatype& st, void* recv, int, const Datatype&, int) const {
    copy(send, recv, scount, st);
  }
  void Allgather(const void* send, int scount, const Datatype& st, void* recv, int, const Datatype&) const {
    copy(send, recv, scount, st);
  }
  void Scatter(const void* send, int scount, const Datatype& st, void* recv, int, const Datatype&, int) const {
    copy(send, recv, scount, st);
  }
  void Send(const void*, int, const Datatype&, int, int) const { single_rank_only("MPI::Comm::Send"); }
  void Recv(void*, int, const Datatype&, int, int) const { single_rank_only("MPI::Comm::Recv"); }
  void Recv(void*, int, const Datatype&, int, int, Status&) const { single_rank_only("MPI::Comm::Recv"); }
  Request Isend(const void*, int, const Datatype&, int, int) const {
    single_rank_only("MPI::Comm::Isend");
    return Request();
  }
  Request Irecv(void*, int, const Datatype&, int, int) const {
    single_rank_only("MPI::Comm::Irecv");
    return Request();
  }
};

class Intracomm : public Comm {
 public:
  Intracomm Dup() const { return Intracomm(); }
  Intracomm Split(int, int) const { return Intracomm(); }
};

static Intracomm COMM_WORLD;

inline void Init() {}
inline void Init(int&, char**&) {}
inline void Finalize() {}
inline bool Is_initialized() { return true; }
inline double Wtime() {
  struct timeval tv;
  gettimeofday(&tv, 0);
  return double(tv.tv_sec) + double(tv.tv_usec) * 1e-6;
}

}  // namespace MPI

#endif
