// Scene-file loader: XML -> plugin objects (src/cpu/parser.cpp semantics).
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

/// Loads a Nori scene description.  Relative file names inside it resolve
/// against the file's directory.  Errors throw NoriException naming the file
/// and the line/column.
NoriObject *loadFromXML(const std::string &filename);

/// The same for a document held in memory; `dir` is where relative file names
/// resolve and `name` is used in error messages.
NoriObject *loadFromXMLString(const std::string &text, const std::string &dir, const std::string &name);

NORI_NAMESPACE_END
