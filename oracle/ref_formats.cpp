// Reference file-format harness: TEST INFRASTRUCTURE ONLY.
//
// Linked by oracle/Makefile into oracle/_ref/libverbref.so. Exposes the
// reference's own Matrix3D/Matrix4D writers and readers (src/Matrix.cpp)
// with plain pointers, so that tests can compare files and arrays with the
// product's format code byte for byte. The reference readers exit the
// process on malformed input, so tests hand them well-formed files only.
#include <cstring>
#include <string>

#include "Matrix.h"

namespace {

Matrix4D<double> m4(const int* s, const double* p) {
    Matrix4D<double> m(s[0], s[1], s[2], s[3]);
    if (p) std::memcpy(m.plane_array, p, sizeof(double) * (size_t)s[0] * s[1] * s[2] * s[3]);
    return m;
}

Matrix3D<double> m3(const int* s, const double* p) {
    Matrix3D<double> m(s[0], s[1], s[2]);
    if (p) std::memcpy(m.plane_array, p, sizeof(double) * (size_t)s[0] * s[1] * s[2]);
    return m;
}

size_t n4(const int* s) { return (size_t)s[0] * s[1] * s[2] * s[3]; }
size_t n3(const int* s) { return (size_t)s[0] * s[1] * s[2]; }

}  // namespace

extern "C" {

// Matrix4D::writeToFile(filename, info) (src/Matrix.cpp:4556-4567)
int ref_plt4_write(const char* path, const int* sizes, const double* data, const char* name, const char* info) {
    Matrix4D<double> m = m4(sizes, data);
    m.name = name ? name : "";
    m.writeToFile(path, info ? info : "");
    return 0;
}

// Matrix4D::writeToFile(filename, grid_w, grid_x, grid_y, grid_z) (:4833-4853)
int ref_plt4_write_grid(const char* path, const int* sizes, const double* gw, const double* gx, const double* gy,
                        const double* gz, const double* data, const char* name) {
    Matrix4D<double> m = m4(sizes, data), w = m4(sizes, gw), x = m4(sizes, gx), y = m4(sizes, gy), z = m4(sizes, gz);
    m.name = name ? name : "";
    m.writeToFile(path, w, x, y, z);
    return 0;
}

// Matrix4D::readFromFile(filename, column) (:4906-4965)
int ref_plt4_read_column(const char* path, int column, const int* sizes, double* out) {
    Matrix4D<double> m = m4(sizes, nullptr);
    m.readFromFile(path, column);
    std::memcpy(out, m.plane_array, sizeof(double) * n4(sizes));
    return 0;
}

// Matrix4D::readFromFile(filename, grid_w, grid_x, grid_y, grid_z) (:5344-5413)
int ref_plt4_read_checked(const char* path, const int* sizes, const double* gw, const double* gx, const double* gy,
                          const double* gz, double* out) {
    Matrix4D<double> m = m4(sizes, nullptr), w = m4(sizes, gw), x = m4(sizes, gx), y = m4(sizes, gy), z = m4(sizes, gz);
    m.readFromFile(path, w, x, y, z);
    std::memcpy(out, m.plane_array, sizeof(double) * n4(sizes));
    return 0;
}

// Matrix4D::writeToBinaryFile / readFromBinaryFile (:5704-5726, :5790-5818)
int ref_pltb4_write(const char* path, const int* sizes, const double* data) {
    m4(sizes, data).writeToBinaryFile(path);
    return 0;
}

int ref_pltb4_read(const char* path, const int* sizes, double* out) {
    Matrix4D<double> m = m4(sizes, nullptr);
    m.readFromBinaryFile(path);
    std::memcpy(out, m.plane_array, sizeof(double) * n4(sizes));
    return 0;
}

// Matrix4D::writeToLstFile(filename, io_method, info, output_folder) (:5762-5784)
int ref_lst_append(const char* filename, int binary, const char* info, const char* output_folder) {
    Matrix4D<double> m;
    m.writeToLstFile(filename, binary ? IOMethod::Binary : IOMethod::ASCII, info, output_folder);
    return 0;
}

// Matrix3D (boundary arrays): writeToFile(filename, info) (:3011-3024),
// writeToFile with grids (:3031-3050), readFromFile with grids
// (:3431-3477), binary (:3781-3838)
int ref_plt3_write(const char* path, const int* sizes, const double* data, const char* name, const char* info) {
    Matrix3D<double> m = m3(sizes, data);
    m.name = name ? name : "";
    m.writeToFile(path, info ? info : "");
    return 0;
}

int ref_plt3_write_grid(const char* path, const int* sizes, const double* gx, const double* gy, const double* gz,
                        const double* data, const char* name) {
    Matrix3D<double> m = m3(sizes, data), x = m3(sizes, gx), y = m3(sizes, gy), z = m3(sizes, gz);
    m.name = name ? name : "";
    m.writeToFile(path, x, y, z);
    return 0;
}

int ref_plt3_read_checked(const char* path, const int* sizes, const double* gx, const double* gy, const double* gz,
                          double* out) {
    Matrix3D<double> m = m3(sizes, nullptr), x = m3(sizes, gx), y = m3(sizes, gy), z = m3(sizes, gz);
    m.readFromFile(path, x, y, z);
    std::memcpy(out, m.plane_array, sizeof(double) * n3(sizes));
    return 0;
}

int ref_pltb3_write(const char* path, const int* sizes, const double* data) {
    m3(sizes, data).writeToBinaryFile(path);
    return 0;
}

int ref_pltb3_read(const char* path, const int* sizes, double* out) {
    Matrix3D<double> m = m3(sizes, nullptr);
    m.readFromBinaryFile(path);
    std::memcpy(out, m.plane_array, sizeof(double) * n3(sizes));
    return 0;
}

}  // extern "C"
