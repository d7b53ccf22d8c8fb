// ScannerInterfileReader.h -- .scan key-file reader (reference
// ScannerInterfileReader.cc:5-52).
#pragma once

#include <ScannerHeader.h>

#include <string>

void readScannerInterfileHeader(const std::string& headerFileName,
                                ScannerHeader& header);
