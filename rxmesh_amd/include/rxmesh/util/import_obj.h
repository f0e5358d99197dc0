// Wavefront OBJ reader for the C++ API (the role of the reference's
// include/rxmesh/util/import_obj.h): 'v' lines become vertices, 'f' lines
// faces (1-based, negative = relative; "a/b/c" keeps the position index);
// polygons are split into a triangle fan.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace rxmesh {

template <typename T>
bool import_obj(const std::string& file_name, std::vector<std::vector<T>>& vertices,
                std::vector<std::vector<uint32_t>>& faces, bool silent = true)
{
    std::ifstream in(file_name);
    if (!in) {
        if (!silent) std::fprintf(stderr, "import_obj: cannot open %s\n", file_name.c_str());
        return false;
    }
    vertices.clear();
    faces.clear();
    std::string line;
    std::vector<long> poly;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string        tag;
        ss >> tag;
        if (tag == "v") {
            std::vector<T> v(3);
            double         x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            v[0] = static_cast<T>(x);
            v[1] = static_cast<T>(y);
            v[2] = static_cast<T>(z);
            vertices.push_back(v);
        } else if (tag == "f") {
            poly.clear();
            std::string tok;
            while (ss >> tok) {
                long i = std::strtol(tok.c_str(), nullptr, 10);
                if (i < 0) i = static_cast<long>(vertices.size()) + i + 1;
                poly.push_back(i - 1);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k)
                faces.push_back({uint32_t(poly[0]), uint32_t(poly[k]), uint32_t(poly[k + 1])});
        }
    }
    return true;
}

}  // namespace rxmesh
