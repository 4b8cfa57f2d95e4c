// PointerSetterBase / PointerSetter<T> (include/PointerSetter.h:25-60): lets a
// communication buffer owner hand an updater the address of its slice.
#ifndef BCM_FACADE_POINTERSETTER_H
#define BCM_FACADE_POINTERSETTER_H

class PointerSetterBase {
 public:
  virtual ~PointerSetterBase() {}
  virtual void setPointer(void* voidPointer) = 0;
};

template <typename T>
class PointerSetter : public PointerSetterBase {
  T** pointer;

 public:
  explicit PointerSetter(T** p) : pointer(p) {}
  void setPointer(void* voidPointer) { *pointer = static_cast<T*>(voidPointer); }
};

#endif
