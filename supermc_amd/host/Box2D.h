// Box2D -- the axis-aligned box of the reference's hot-spot code
// (Box2D.h:16-74, Box2D.cpp), for callers of MCnucl::getHotSpots.  The
// public interface and the arithmetic are the reference's, because box
// edges feed the integer cell offsets of the eccentricity mask
// (MakeDensity.cpp:2341-2359): the centre is the mean of the edges, moving
// the centre shifts every edge by the same difference.
//
// Difference: overUnion prints nothing (the reference's check after the
// min/max can only fire on NaN edges).
#ifndef SMC_HOST_BOX2D_H
#define SMC_HOST_BOX2D_H

#include <algorithm>
#include <iostream>
#include <sstream>
#include <string>

class Box2D {
 public:
  Box2D() = default;
  Box2D(double xl, double xr, double yl, double yr)
      : xl_(xl), xr_(xr), yl_(yl), yr_(yr), xc_((xr + xl) / 2.0), yc_((yr + yl) / 2.0) {}

  double getXL() const { return xl_; }
  double getXR() const { return xr_; }
  double getYL() const { return yl_; }
  double getYR() const { return yr_; }
  double getX() const { return xc_; }
  double getY() const { return yc_; }

  // moves the box: every edge shifts by the change of the centre
  void setCenter(double x, double y) {
    const double sx = x - xc_, sy = y - yc_;
    xl_ += sx;
    xr_ += sx;
    yl_ += sy;
    yr_ += sy;
    xc_ = x;
    yc_ = y;
  }
  void setX(double x) { setCenter(x, yc_); }
  void setY(double y) { setCenter(xc_, y); }
  // edges at the centre -+ half the lengths
  void setDimensions(double lx, double ly) {
    xl_ = xc_ - lx / 2;
    xr_ = xc_ + lx / 2;
    yl_ = yc_ - ly / 2;
    yr_ = yc_ + ly / 2;
  }
  void setSquareDimensions(double size) { setDimensions(size, size); }

  // the overlap of two boxes (not positive when they do not meet)
  Box2D intersection(const Box2D& o) const {
    return Box2D(std::max(xl_, o.xl_), std::min(xr_, o.xr_), std::max(yl_, o.yl_), std::min(yr_, o.yr_));
  }
  // grows this box to the smallest box holding both
  void overUnion(const Box2D& o) {
    xl_ = std::min(o.xl_, xl_);
    xr_ = std::max(o.xr_, xr_);
    yl_ = std::min(o.yl_, yl_);
    yr_ = std::max(o.yr_, yr_);
    xc_ = (xl_ + xr_) / 2.0;
    yc_ = (yl_ + yr_) / 2.0;
  }
  bool isPositive() const { return xl_ < xr_ && yl_ < yr_; }
  double getArea() const { return (xr_ - xl_) * (yr_ - yl_); }
  std::string toString() const {
    std::ostringstream os;
    os << xl_ << " " << xr_ << " " << yl_ << " " << yr_;
    return os.str();
  }
  void printToConsole() const {
    std::cout << "(" << xl_ << "," << xr_ << "," << yl_ << "," << yr_ << ")" << std::endl;
  }

 private:
  double xl_ = 0, xr_ = 0, yl_ = 0, yr_ = 0;
  double xc_ = 0, yc_ = 0;
};

#endif
