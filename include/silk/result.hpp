// Status type returned by every silk::World call.
//
// Declaration-identical to the reference interface
// (silk/include/silk/result.hpp:11-80) so that callers of silk compile
// unchanged against this library.  Success carries no detail; an error carries
// an ErrorCode and an optional human readable reason.
#pragma once

#include <string>
#include <string_view>

namespace silk {

/// Error categories.  The enumerator order is part of the ABI and matches the
/// reference (result.hpp:22-35).  NoCudaSupport keeps its name: on this
/// library it means "no usable HIP device".
enum class ErrorCode {
  Unknown,
  NotInitialized,
  InvalidConfig,
  TooManyBody,
  InvalidHandle,
  InvalidMesh,
  InvalidPin,
  InvalidPosition,
  CholeskyDecompositionFail,
  NeedInitSolverFirst,
  IterativeSolveFail,
  NoCudaSupport
};

/// Non-throwing result value.  `operator bool` is true on success.
class Result {
 public:
  bool success = true;
  ErrorCode code = ErrorCode::Unknown;
  std::string detail;

 public:
  /// A success value.
  static Result ok();

  /// An error value with a category and optional reason.
  static Result error(ErrorCode error_code, std::string_view detail = {});

  /// True on success.
  operator bool() const;

  /// "OK", "<Code>" or "<Code>. Reason: <detail>.".
  std::string to_string() const;

 private:
  Result() = default;
};

}  // namespace silk
