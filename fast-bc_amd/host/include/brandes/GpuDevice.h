/* The device used by the GPU-backed computers: device 0 by default, or one
 * chosen with setDevice() before the first computation. */
#ifndef FASTBC_AMD_GPUDEVICE_H
#define FASTBC_AMD_GPUDEVICE_H

#include "../../fbc_host.h"

#include <memory>
#include <mutex>

namespace fastbc {
namespace brandes {

inline fbchost::Device &gpuDevice(int select = -1)
{
    static std::mutex mu;
    static std::unique_ptr<fbchost::Device> dev;
    std::lock_guard<std::mutex> lk(mu);
    if (!dev) dev.reset(new fbchost::Device(select < 0 ? 0 : select));
    return *dev;
}

/* FBC_PREC_F64 (the reference's arithmetic) unless set otherwise. */
inline int &gpuPrecision()
{
    static int prec = FBC_PREC_F64;
    return prec;
}

} // namespace brandes
} // namespace fastbc

#endif
