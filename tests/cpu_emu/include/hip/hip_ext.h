// CPU emulation of hipExtLaunchKernelGGL: the events are recorded around the
// launch, in stream order.
#pragma once
#include <hip/hip_runtime.h>

template <typename... KArgs, typename... Args>
inline void hipExtLaunchKernelGGL(void (*kernel)(KArgs...), const dim3& grid, const dim3& block,
                                  unsigned shmem, hipStream_t stream, hipEvent_t start,
                                  hipEvent_t stop, unsigned flags, Args&&... args) {
  (void)shmem;
  (void)flags;
  if (start) hipEventRecord(start, stream);
  ::hipemu::launch(stream, kernel, grid, block, std::forward<Args>(args)...);
  if (stop) hipEventRecord(stop, stream);
}
