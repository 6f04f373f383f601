// FileIO.h -- the reference's binary outputs and restart files
// (src/FileIO.mm), written from the device state through the C-ABI.
//
//   bin/<ProjectName>_EMField_%08d.bin            outputField     FileIO.mm:41-132
//   bin/<ProjectName>_Moments_<species>_%08d.bin  outputMoments   FileIO.mm:134-226
//   bin/<ProjectName>_PhaseSpace_<species>_<i>.bin outputPhaseSpace FileIO.mm:4-39 (appended)
//   bin/<ProjectName>_ProgressData_<species>_%08d.bin saveProgress FileIO.mm:228-287
//
// A field or moment file is a header (ngx, ngy, ngb as int32; dx, dy as
// float32) followed by records of a 32-byte name, an int32 type id (0 node,
// 1 x-shifted, 2 y-shifted, 3 both, 4 potential) and the FP32 array in SI
// units, as tools/field.py reads them.  A progress file is the particle
// array as 24-byte records {x, y, vx, vy, vz: float32; flag: int32}.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "Init.h"
#include "Modules.h"

namespace gample {

// the output directory of every file above ("bin", created when missing)
void ensureOutputDir();
// each returns its log section: min and max of the written arrays in the
// written units (outputField FileIO.mm:68-101, outputMoments :162-197)
std::map<std::string, std::string> outputField(int cycle, EMField &fld, const Init &init);
std::map<std::string, std::string> outputMoments(int cycle, EMField &fld, const std::vector<double> &mom,
                                                 const std::string &pName, const Init &init);
void outputPhaseSpace(int cycle, Particle &p, const Init &init);
void saveProgress(int cycle, std::vector<Particle> &ptcl, const Init &init);
// reads bin/<RestartName>_ProgressData_<species>_%08d.bin into each species
// and shifts positions by PtclPosOffset (FileIO.mm:289-352)
void loadProgress(int cycle, std::vector<Particle> &ptcl, const Init &init);

}  // namespace gample
