// kp_error.h -- thread-local last-error message behind kp_last_error().
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/kpolrad.h"

namespace kp {

void set_error(const std::string &msg);
void clear_error();

// Error carrying a kp_status; thrown inside the library, caught at the ABI.
struct error : std::runtime_error {
    kp_status status;
    error(kp_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// Runs f() and converts any exception into a status + message.
template <class F> kp_status guard(F &&f)
{
    try {
        clear_error();
        f();
        return KP_OK;
    } catch (const error &e) {
        set_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &e) {
        set_error(std::string("allocation failed: ") + e.what());
        return KP_ERR_ALLOC;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KP_ERR_ARG;
    } catch (...) {
        set_error("unknown error");
        return KP_ERR_ARG;
    }
}

} // namespace kp
