///
/// @file   primesieve_error.hpp
/// @brief  Exception type of the count path.  Same name, base class and
///         constructor as the reference's own header
///         (include/primesieve/primesieve_error.hpp:18-28), so that code
///         written against the reference catches it unchanged.  When the
///         backend is built inside the reference tree, that header is used
///         instead of this one.
///
#ifndef PRIMESIEVE_ERROR_HPP
#define PRIMESIEVE_ERROR_HPP

#include <stdexcept>
#include <string>

namespace primesieve {

class primesieve_error : public std::runtime_error
{
public:
  primesieve_error(const std::string& msg)
    : std::runtime_error(msg)
  { }
};

} // namespace primesieve

#endif
