// Stand-in for senk_amd/csrc/internal.hpp when the host-only MatrixMarket reader
// (host_io.cpp) is built alone with g++ under the address and undefined-behaviour
// sanitizers (tests/test_host_sanitizers.py): the status codes and the error hook.
#pragma once
#include <string>

#include "senk_amd.h"

namespace senk_amd {
void record_error(const std::string &msg);
}
