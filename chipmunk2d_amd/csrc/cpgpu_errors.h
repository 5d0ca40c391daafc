// cpgpu_errors.h -- the message of the last failed call, shared by the library's builds.
#pragma once

#include <string>

extern thread_local std::string cpgpu_last_error;
