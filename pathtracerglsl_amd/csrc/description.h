// description.h — the ProgramDescription lowering (description.cpp), used by the C ABI.
#pragma once
#include <cstddef>
#include <string>

#include "../../include/ptmi.h"

pt_status ptmi_lower_description(const char* json, size_t len, int32_t width, int32_t height, int32_t frames_in_flight,
                                 pt_program* out, std::string& err);
