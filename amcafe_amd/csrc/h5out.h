// h5out.h -- VoxelsCA::WriteToHDF1 (S/VoxelsCA.C:1592-1856) on host buffers.
#pragma once
#include <string>

namespace amcafe {
// 9 floats per grain: RGB of the IPF colour of z (build direction), x and y
void ipf_colours(const double* cTheta, int nGrain, float* rgb9);
// the reference's output file; 0 or -1 with err set
int write_hdf5(const char* path, const int nX[3], const double dX[3], const int* gID, const int* vState,
               const int* Build, const double* TempCurr, const double* cTheta, int nGrain, std::string& err);
}  // namespace amcafe
