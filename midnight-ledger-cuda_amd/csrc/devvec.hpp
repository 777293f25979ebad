// Resident Fr vector behind the reference's opaque device-vector handle
// (scalar_device_vec_t, bls12_381_g1_msm.cu:1110-1114): its length, the
// lane that created it, and the device allocation.
#pragma once
#include "fields.hpp"
#include "host_fr.hpp"
#include "runtime.hpp"

namespace mh {

struct DeviceVec {
    uint32_t lane = 0;
    size_t len = 0;
    DevBuf<Fr> values;
};

Fr to_fr(const RawFr &r);
HostFr to_hfr(const RawFr &r);

}  // namespace mh
