// Minimal OpenMM-compatible core: the exception type of the OpenMM API.
#pragma once
#include <exception>
#include <string>

namespace OpenMM {

class OpenMMException : public std::exception {
public:
    explicit OpenMMException(const std::string& message) : message(message) {}
    ~OpenMMException() throw() override {}
    const char* what() const throw() override { return message.c_str(); }

private:
    std::string message;
};

}  // namespace OpenMM
