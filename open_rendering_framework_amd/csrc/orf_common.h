// orf_common.h -- status codes, last-error text and HIP error checks for the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "../../include/orf_rt.h"

namespace orf {

void set_last_error(const std::string& msg);

struct Error {
    orf_status code;
    std::string msg;
};

}  // namespace orf

#define ORF_THROW(code_, msg_) throw ::orf::Error{(code_), (msg_)}

#define ORF_HIP_CHECK(expr)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            ORF_THROW(ORF_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define ORF_REQUIRE(cond, msg_)                                   \
    do {                                                          \
        if (!(cond)) ORF_THROW(ORF_ERROR_INVALID_VALUE, (msg_));  \
    } while (0)

// Wraps a C-ABI body: exceptions become status codes + orf_last_error().
#define ORF_API_BEGIN try {
#define ORF_API_END                                                          \
    }                                                                        \
    catch (const ::orf::Error& e) {                                          \
        ::orf::set_last_error(e.msg);                                        \
        return e.code;                                                       \
    }                                                                        \
    catch (const std::bad_alloc&) {                                          \
        ::orf::set_last_error("host allocation failed");                    \
        return ORF_ERROR_OUT_OF_MEMORY;                                      \
    }                                                                        \
    catch (const std::exception& e) {                                        \
        ::orf::set_last_error(e.what());                                     \
        return ORF_ERROR_UNKNOWN;                                            \
    }                                                                        \
    return ORF_SUCCESS;
