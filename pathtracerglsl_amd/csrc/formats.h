// formats.h — on-disk scene inputs (formats.cpp), used by the C ABI: the Wavefront OBJ
// reader and the `.inst` instance-cloud reader.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/ptmi.h"

struct pt_mesh_data {
    std::vector<float> positions, normals, uvs;   // 3 / 3 / 2 floats per output vertex
    std::vector<uint32_t> indices;                // 3 per triangle
    bool has_normals = false, has_uvs = false;
};

pt_status ptmi_parse_obj(const char* text, size_t len, uint32_t flags, pt_mesh_data& out, std::string& err);
pt_status ptmi_parse_inst(const uint8_t* buf, size_t len, pt_inst_table& out, std::string& err);
