// Host runtime helpers shared by the library's translation units: error
// types and the status-code guard of the C ABI, device selection, growable
// device buffers and events.
#pragma once
#include "../padenti_hip.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace pd {
namespace engine {

inline thread_local std::string g_err;

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NoMem : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ParityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

template <class Fn>
inline int guarded(Fn &&fn)
{
  try {
    fn();
    return PD_OK;
  } catch (const ArgError &e) {
    g_err = e.what();
    return PD_ERR_ARG;
  } catch (const NoMem &e) {
    g_err = e.what();
    return PD_ERR_NOMEM;
  } catch (const HipError &e) {
    g_err = e.what();
    return PD_ERR_HIP;
  } catch (const StateError &e) {
    g_err = e.what();
    return PD_ERR_STATE;
  } catch (const ParityError &e) {
    g_err = e.what();
    return PD_ERR_PARITY;
  } catch (const std::exception &e) {
    g_err = e.what();
    return PD_ERR_STATE;
  }
}

inline void require(bool ok, const std::string &msg)
{
  if (!ok) throw ArgError(msg);
}

inline void set_device(int device)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw HipError("no HIP device available");
  require(device >= 0 && device < n, "device index out of range");
  HIPCHK(hipSetDevice(device));
}

// Growable device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  void ensure(size_t bytes)
  {
    if (bytes <= cap && p) return;
    release();
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw NoMem("hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    cap = bytes;
  }
  template <class T>
  T *as() const
  {
    return (T *)p;
  }
};

struct Event {
  hipEvent_t e = nullptr;
  Event() { HIPCHK(hipEventCreate(&e)); }
  ~Event()
  {
    if (e) (void)hipEventDestroy(e);
  }
};

inline double elapsed_ms(const Event &a, const Event &b)
{
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a.e, b.e));
  return (double)ms;
}

}  // namespace engine
}  // namespace pd
