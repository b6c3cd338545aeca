// ExceptionT.h -- error codes and the throwing convention of the host layer.
//
// The codes and their meaning are the reference's ExceptionT::CodeT
// (toolbox/src/common/ExceptionT.h:17-30); the C ABI of libtahoe_gpu returns
// the same integers, so a non-zero status becomes the exception the
// reference would have thrown at that point.
#ifndef TAHOE_HIP_EXCEPTION_T_H
#define TAHOE_HIP_EXCEPTION_T_H

#include <stdexcept>
#include <string>

#include "tahoe_gpu.h"

namespace tahoe_hip {

class ExceptionT : public std::runtime_error {
public:
    enum CodeT {
        kNoError = 0,
        kGeneralFail = 1,
        kStop = 2,
        kOutOfMemory = 3,
        kOutOfRange = 4,
        kSizeMismatch = 5,
        kBadInputValue = 6,
        kBadJacobianDet = 7,
        kMPIFail = 8,
        kDatabaseFail = 9,
        kBadHeartBeat = 10,
        kTypeMismatch = 11
    };

    ExceptionT(CodeT code, const std::string& caller, const std::string& msg)
        : std::runtime_error(std::string(ToString(code)) + " in " + caller + (msg.empty() ? "" : ": " + msg)),
          fCode(code)
    {
    }
    CodeT Code() const { return fCode; }

    static const char* ToString(CodeT code)
    {
        switch (code) {
        case kNoError: return "no error";
        case kGeneralFail: return "general fail";
        case kStop: return "stop";
        case kOutOfMemory: return "out of memory";
        case kOutOfRange: return "out of range";
        case kSizeMismatch: return "size mismatch";
        case kBadInputValue: return "bad input value";
        case kBadJacobianDet: return "bad jacobian determinant";
        case kMPIFail: return "MPI message passing failure";
        case kDatabaseFail: return "database read failure";
        case kBadHeartBeat: return "bad heartbeat";
        case kTypeMismatch: return "type mismatch";
        }
        return "unknown";
    }

    [[noreturn]] static void Throw(CodeT code, const std::string& caller, const std::string& msg = "")
    {
        throw ExceptionT(code, caller, msg);
    }
    [[noreturn]] static void GeneralFail(const std::string& caller, const std::string& msg = "")
    {
        Throw(kGeneralFail, caller, msg);
    }
    [[noreturn]] static void BadInputValue(const std::string& caller, const std::string& msg = "")
    {
        Throw(kBadInputValue, caller, msg);
    }
    [[noreturn]] static void SizeMismatch(const std::string& caller, const std::string& msg = "")
    {
        Throw(kSizeMismatch, caller, msg);
    }
    [[noreturn]] static void OutOfRange(const std::string& caller, const std::string& msg = "")
    {
        Throw(kOutOfRange, caller, msg);
    }

private:
    CodeT fCode;
};

// status of a C ABI call -> exception carrying the handle's message
inline void tg_check(int status, const char* caller, const tg_ctx* ctx = nullptr)
{
    if (status != TG_OK)
        ExceptionT::Throw(ExceptionT::CodeT(status), caller, ctx ? tg_last_error(ctx) : "device handle not created");
}

}  // namespace tahoe_hip

#endif
