// ReconstructorAbstractFactory.h -- Include/ReconstructorAbstractFactory.h:5-12
// (the class name keeps the reference's spelling).
#pragma once
#include <string>

class Reconstructor;

class ReconstrctorAbstractFactory
{
public:
    ReconstrctorAbstractFactory() {}
    ~ReconstrctorAbstractFactory() {}
    Reconstructor *createReconstructor(const std::string &str);
};
