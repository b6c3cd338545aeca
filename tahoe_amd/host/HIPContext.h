// HIPContext.h -- owner of one device problem (a tg_ctx handle).
#ifndef TAHOE_HIP_HIP_CONTEXT_H
#define TAHOE_HIP_HIP_CONTEXT_H

#include "ExceptionT.h"

namespace tahoe_hip {

class HIPContext {
public:
    explicit HIPContext(int device = 0) { tg_check(tg_create(&fCtx, device), "HIPContext"); }
    ~HIPContext()
    {
        if (fCtx) tg_destroy(fCtx);
    }
    HIPContext(const HIPContext&) = delete;
    HIPContext& operator=(const HIPContext&) = delete;

    tg_ctx* Handle() const { return fCtx; }
    // status of a call on this handle -> ExceptionT with the handle's message
    void Check(int status, const char* caller) const { tg_check(status, caller, fCtx); }

private:
    tg_ctx* fCtx = nullptr;
};

}  // namespace tahoe_hip

#endif
